"""Every public function of the oracle names the check that pins it (oracle.dga_oracle.PARITY)."""
import inspect

from oracle import dga_oracle as O


def test_every_oracle_function_has_a_parity_entry():
    funcs = [n for n, f in inspect.getmembers(O, inspect.isfunction) if f.__module__ == O.__name__ and not n.startswith("_")]
    missing = [f for f in funcs if f not in O.PARITY]
    assert not missing, missing
    assert all(isinstance(v, str) and v for v in O.PARITY.values())
    stale = [k for k in O.PARITY if k not in funcs]
    assert not stale, stale
