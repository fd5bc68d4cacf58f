"""Read-only access to HDF5 files in the layout h5py writes by default, without h5py.

The DMFT input of the reference (w2dynamics output, read with h5py in
dmft_interface.py:125-262) is an HDF5 file. So that the package needs no h5py, this module
parses the file format itself, for the structures such files use (those of the reference's
end_2_end/1p-data.hdf5): superblock version 0, groups as symbol tables (version-1 B-trees of
symbol-table nodes with a local heap of names, soft links), version-1 object headers with their
continuation blocks, dataspace messages of version 1, contiguous storage (layout message version
3), and datatypes of fixed-point, floating-point, compound, enumeration and variable-length
string class, in datasets and in attribute messages of version 1 (variable-length strings from
global heaps). Numeric data comes back as numpy arrays (a compound of two equal floats named
r/i or re/im as complex), strings as str. Nothing in the file is executed: every value is
decoded from its declared type. Other structures (chunked or filtered storage, version-2
object headers, dense attribute or link storage, newer message versions) raise
NotImplementedError.
"""
import numpy as np

_SIGNATURE = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class _Reader:
    def __init__(self, data):
        self.d = data

    def u(self, off, n):
        return int.from_bytes(self.d[off:off + n], "little")


class Datatype:
    """A decoded datatype: kind "num" with its numpy dtype, or kind "vlen-string"."""

    def __init__(self, kind, size, dtype=None):
        self.kind, self.size, self.dtype = kind, size, dtype


class File:
    """An HDF5 file opened for reading: f["group/dataset"] returns a Group or the dataset's
    value as a numpy array; Group.attrs maps attribute names to values."""

    def __init__(self, path):
        with open(path, "rb") as fh:
            self._d = fh.read()
        self._r = _Reader(self._d)
        base = None
        for off in (0, 512, 1024, 2048, 4096, 8192):
            if self._d[off:off + 8] == _SIGNATURE:
                base = off
                break
        if base is None:
            raise ValueError(f"{path}: not an HDF5 file")
        ver = self._d[base + 8]
        if ver != 0:
            raise NotImplementedError(f"HDF5 superblock version {ver} is not supported")
        self.so, self.sl = self._d[base + 13], self._d[base + 14]
        if self.so != 8 or self.sl != 8:
            raise NotImplementedError("only 8-byte offsets and lengths are supported")
        p = base + 24
        self.base_addr = self._r.u(p, 8)
        root_entry = p + 4 * self.so
        self.root = Group(self, self._r.u(root_entry + self.so, 8), "/")

    def __getitem__(self, path):
        return self.root[path]

    def keys(self):
        return self.root.keys()

    # --- low-level helpers -------------------------------------------------------------
    def addr(self, a):
        return self.base_addr + a

    def messages(self, ohdr):
        """[(type, flags, offset, size)] of the object header at file address ohdr."""
        d, r = self._d, self._r
        a = self.addr(ohdr)
        if d[a:a + 4] == b"OHDR":
            raise NotImplementedError("version-2 object headers are not supported")
        if d[a] != 1:
            raise ValueError(f"bad object header version {d[a]} at {ohdr}")
        nmsg = r.u(a + 2, 2)
        size = r.u(a + 8, 4)
        blocks = [(a + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(out) < nmsg:
                mtype, msize, flags = r.u(p, 2), r.u(p + 2, 2), d[p + 4]
                body = p + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self.addr(r.u(body, 8)), r.u(body + 8, 8)))
                out.append((mtype, flags, body, msize))
                p = body + msize
        return out

    def local_heap_name(self, heap, off):
        d, r = self._d, self._r
        a = self.addr(heap)
        if d[a:a + 4] != b"HEAP":
            raise ValueError("bad local heap")
        data = self.addr(r.u(a + 24, 8))
        end = d.index(b"\x00", data + off)
        return d[data + off:end].decode("utf-8")

    def group_entries(self, btree, heap):
        """{name: object header address, or the target path (str) of a soft link} of a
        symbol-table group."""
        d, r = self._d, self._r
        out = {}
        stack = [btree]
        while stack:
            a = self.addr(stack.pop())
            if d[a:a + 4] != b"TREE" or d[a + 4] != 0:
                raise ValueError("bad group B-tree node")
            level, used = d[a + 5], r.u(a + 6, 2)
            p = a + 8 + 2 * self.so
            children = []
            for i in range(used):
                p += self.sl  # key
                children.append(r.u(p, 8))
                p += self.so
            if level > 0:
                stack.extend(children)
                continue
            for snod in children:
                s = self.addr(snod)
                if d[s:s + 4] != b"SNOD":
                    raise ValueError("bad symbol table node")
                nsym = r.u(s + 6, 2)
                e = s + 8
                for _ in range(nsym):
                    name = self.local_heap_name(heap, r.u(e, 8))
                    if r.u(e + 16, 4) == 2:  # soft link: the target path is in the local heap
                        out[name] = self.local_heap_name(heap, r.u(e + 24, 4))
                    else:
                        out[name] = r.u(e + 8, 8)
                    e += 2 * self.so + 24
        return out

    def global_heap_object(self, coll, index):
        d, r = self._d, self._r
        a = self.addr(coll)
        if d[a:a + 4] != b"GCOL":
            raise ValueError("bad global heap collection")
        size = r.u(a + 8, 8)
        p, end = a + 16, a + size
        while p + 16 <= end:
            idx, osize = r.u(p, 2), r.u(p + 8, 8)
            if idx == 0:
                break
            if idx == index:
                return d[p + 16:p + 16 + osize]
            p += 16 + ((osize + 7) // 8) * 8
        raise KeyError(f"global heap object {index} not found")

    # --- messages ----------------------------------------------------------------------
    def dataspace(self, p):
        d, r = self._d, self._r
        ver, rank = d[p], d[p + 1]
        if ver != 1:
            raise NotImplementedError(f"dataspace version {ver}")
        return tuple(r.u(p + 8 + 8 * i, 8) for i in range(rank))  # rank 0: a scalar

    def datatype(self, p):
        d, r = self._d, self._r
        cls, ver = d[p] & 0x0F, d[p] >> 4
        if ver != 1:
            raise NotImplementedError(f"datatype message version {ver}")
        bits = r.u(p + 1, 3)
        size = r.u(p + 4, 4)
        q = p + 8
        if cls == 0:  # fixed point
            order = ">" if bits & 1 else "<"
            kind = "i" if bits & 8 else "u"
            return Datatype("num", size, np.dtype(f"{order}{kind}{size}"))
        if cls == 1:  # floating point
            order = ">" if bits & 1 else "<"
            return Datatype("num", size, np.dtype(f"{order}f{size}"))
        if cls == 6:  # compound (version 1: names padded to 8 bytes, 4-byte offsets, a dimension block)
            members = []
            for _ in range(bits & 0xFFFF):
                end = d.index(b"\x00", q)
                name = d[q:end].decode("utf-8")
                q += ((end - q) // 8 + 1) * 8
                off = r.u(q, 4)
                q += 4 + 1 + 3 + 4 + 4 + 16
                sub = self.datatype(q)
                q += self._datatype_length(q)
                members.append((name, off, sub))
            names = [m[0] for m in members]
            if (len(members) == 2 and all(m[2].dtype.kind == "f" for m in members)
                    and members[0][2].dtype == members[1][2].dtype and names in (["r", "i"], ["re", "im"])
                    and members[0][1] == 0 and members[1][1] == members[0][2].size and size == 2 * members[0][2].size):
                fs = members[0][2].size  # complex, as h5py stores it
                return Datatype("num", size, np.dtype(f"{members[0][2].dtype.byteorder.replace('=', '<')}c{2 * fs}"))
            dt = np.dtype({"names": names, "formats": [m[2].dtype for m in members],
                           "offsets": [m[1] for m in members], "itemsize": size})
            return Datatype("num", size, dt)
        if cls == 8:  # enumeration (h5py's booleans): the values of its integer base type
            return Datatype("num", size, self.datatype(q).dtype)
        if cls == 9 and (bits & 0x0F) == 1:  # variable-length string
            return Datatype("vlen-string", size)
        raise NotImplementedError(f"HDF5 datatype class {cls}")

    def _datatype_length(self, p):
        """Bytes of the (fixed- or floating-point) member datatype at p."""
        cls = self._d[p] & 0x0F
        if cls == 0:
            return 8 + 4  # bit offset, bit precision
        if cls == 1:
            return 8 + 12  # bit offset, precision, exponent and mantissa fields, exponent bias
        raise NotImplementedError(f"compound member of datatype class {cls}")

    def decode(self, raw, dt, shape):
        """Values of `shape` elements of datatype dt from the bytes raw."""
        count = int(np.prod(shape)) if shape else 1
        if dt.kind == "num":
            arr = np.frombuffer(raw[:count * dt.size], dtype=dt.dtype).copy()
            if arr.dtype.byteorder == ">":
                arr = arr.astype(arr.dtype.newbyteorder("="))
            return arr.reshape(shape) if shape else arr[0]
        vals = []  # variable-length strings: length, then the global heap collection and object index
        for i in range(count):
            p = i * 16
            n = int.from_bytes(raw[p:p + 4], "little")
            coll = int.from_bytes(raw[p + 4:p + 12], "little")
            idx = int.from_bytes(raw[p + 12:p + 16], "little")
            vals.append(self.global_heap_object(coll, idx)[:n].decode("utf-8") if coll else "")
        return np.array(vals, dtype=object).reshape(shape) if shape else vals[0]

    def attributes(self, ohdr):
        d, r = self._d, self._r
        out = {}
        pad = lambda n: ((n + 7) // 8) * 8  # noqa: E731  (version-1 fields are 8-byte aligned)
        for mtype, _, p, _ in self.messages(ohdr):
            if mtype == 0x15:
                raise NotImplementedError("dense attribute storage is not supported")
            if mtype != 0x0C:
                continue
            if d[p] != 1:
                raise NotImplementedError(f"attribute message version {d[p]}")
            nsz, tsz, ssz = r.u(p + 2, 2), r.u(p + 4, 2), r.u(p + 6, 2)
            q = p + 8
            name = d[q:q + nsz].split(b"\x00")[0].decode("utf-8")
            q += pad(nsz)
            dt = self.datatype(q)
            q += pad(tsz)
            shape = self.dataspace(q)
            q += pad(ssz)
            count = int(np.prod(shape)) if shape else 1
            out[name] = self.decode(d[q:q + count * dt.size], dt, shape)
        return out

    def dataset_value(self, ohdr):
        d, r = self._d, self._r
        shape = dt = layout = None
        for mtype, _, p, _ in self.messages(ohdr):
            if mtype == 0x01:
                shape = self.dataspace(p)
            elif mtype == 0x03:
                dt = self.datatype(p)
            elif mtype == 0x08:
                layout = p
            elif mtype == 0x0B:
                raise NotImplementedError("filtered (compressed) datasets are not supported")
        if d[layout] != 3 or d[layout + 1] != 1:
            raise NotImplementedError("only contiguous storage (layout message version 3) is supported")
        a, n = r.u(layout + 2, 8), r.u(layout + 10, 8)
        count = int(np.prod(shape)) if shape else 1
        # storage never written (undefined address) holds the fill value 0
        raw = bytes(count * dt.size) if a == _UNDEF else d[self.addr(a):self.addr(a) + n]
        return self.decode(raw, dt, shape)


class Group:
    """A group of an HDF5 file: g[name] (a path relative to it), g.keys(), g.attrs."""

    def __init__(self, f, ohdr, name):
        self.f, self.ohdr, self.name = f, ohdr, name
        btree = heap = None
        for mtype, _, p, _ in f.messages(ohdr):
            if mtype == 0x11:
                btree, heap = f._r.u(p, 8), f._r.u(p + 8, 8)
            elif mtype in (0x02, 0x06, 0x0A):
                raise NotImplementedError("groups with link messages (new-style groups) are not supported")
        self._entries = f.group_entries(btree, heap) if btree is not None else {}
        self._attrs = None

    @property
    def attrs(self):
        if self._attrs is None:
            self._attrs = self.f.attributes(self.ohdr)
        return self._attrs

    def keys(self):
        return list(self._entries)

    def __contains__(self, path):
        try:
            self[path]
        except KeyError:
            return False
        return True

    def __iter__(self):
        return iter(self._entries)

    def __getitem__(self, path):
        return self._get(path, 0)

    def _get(self, path, depth):
        if depth > 16:
            raise KeyError(f"{path}: too many soft links")
        node = self.f.root if path.startswith("/") else self
        for part in [p for p in path.split("/") if p]:
            if not isinstance(node, Group):
                raise KeyError(path)
            if part not in node._entries:
                raise KeyError(f"{path}: '{part}' not found")
            target = node._entries[part]
            if isinstance(target, str):  # soft link, relative to the group that holds it
                node = node._get(target, depth + 1)
                continue
            types = {m[0] for m in self.f.messages(target)}
            node = Group(self.f, target, part) if 0x11 in types else Dataset(self.f, target, part)
        return node


class Dataset:
    """A dataset: ds[()] is its value (numpy array or scalar), ds.attrs its attributes."""

    def __init__(self, f, ohdr, name):
        self.f, self.ohdr, self.name = f, ohdr, name

    @property
    def attrs(self):
        return self.f.attributes(self.ohdr)

    def __getitem__(self, key):
        val = self.f.dataset_value(self.ohdr)
        if key == () or key is Ellipsis:
            return val
        return val[key]

    @property
    def shape(self):
        for mtype, _, p, _ in self.f.messages(self.ohdr):
            if mtype == 0x01:
                return self.f.dataspace(p)
        return None
