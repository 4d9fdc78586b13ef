"""Small utilities (quantumflow/utils.py): map inversion, bit lists and
graph6 conversion."""

from typing import Sequence

import networkx as nx

__all__ = ['invert_map', 'bitlist_to_int', 'int_to_bitlist', 'from_graph6',
           'to_graph6']


def invert_map(mapping: dict, one_to_one: bool = True) -> dict:
    """Invert a dict; with one_to_one=False values map to sets of keys
    (utils.py:38-49)."""
    if one_to_one:
        return {value: key for key, value in mapping.items()}
    inv = {}
    for key, value in mapping.items():
        inv.setdefault(value, set()).add(key)
    return inv


def bitlist_to_int(bitlist: Sequence[int]) -> int:
    """[1, 0, 0] -> 4 (utils.py:52-59)."""
    return int(''.join([str(d) for d in bitlist]), 2)


def int_to_bitlist(x: int, pad: int = None) -> Sequence[int]:
    """4 -> [1, 0, 0], left-padded with zeros to `pad` bits (utils.py:62-75)."""
    if pad is None:
        form = '{:0b}'
    else:
        form = '{:0' + str(pad) + 'b}'
    return [int(b) for b in form.format(x)]


def from_graph6(graph6: str) -> nx.Graph:
    """graph6 text -> networkx graph (utils.py:97-100)."""
    return nx.from_graph6_bytes(bytes(graph6, "utf-8"))


def to_graph6(graph: nx.Graph) -> str:
    """networkx graph -> graph6 text (utils.py:103-106)."""
    return nx.to_graph6_bytes(graph)[10:-1].decode("utf-8")
