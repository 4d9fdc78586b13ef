"""Directed-acyclic-graph view of a circuit (quantumflow/dagcircuit.py:25-170).

Nodes are the operations plus an ('in', q) and an ('out', q) node per qubit;
each wire (qubit) links consecutive operations on it.  Iteration yields the
operations in a topological order (their circuit order).  Implemented with
plain adjacency maps (the reference builds a networkx MultiDiGraph), which
also keeps ``components`` working on current networkx releases.
"""

from typing import Dict, Iterable, Iterator, List

from .circuits import Circuit
from .ops import Gate, Operation
from .qubits import Qubit, Qubits
from .states import State

__all__ = ['DAGCircuit']


class DAGCircuit(Operation):
    """A circuit as a DAG of operations linked along qubit wires."""

    def __init__(self, elements: Iterable[Operation]) -> None:
        self._ops: List[Operation] = []
        # per node: {qubit: successor node} / {qubit: predecessor node}
        self._next: Dict[int, Dict[Qubit, object]] = {}
        self._prev: Dict[int, Dict[Qubit, object]] = {}
        last: Dict[Qubit, object] = {}
        self._wires: List[Qubit] = []
        for elem in elements:
            if isinstance(elem, tuple):        # in/out nodes of another DAG
                continue
            self._ops.append(elem)
            self._prev[id(elem)] = {}
            self._next[id(elem)] = {}
            for q in elem.qubits:
                if q not in last:
                    last[q] = ('in', q)
                    self._wires.append(q)
                p = last[q]
                self._prev[id(elem)][q] = p
                if isinstance(p, Operation):
                    self._next[id(p)][q] = elem
                last[q] = elem
        for q, p in last.items():
            if isinstance(p, Operation):
                self._next[id(p)][q] = ('out', q)

    @property
    def qubits(self) -> Qubits:
        return tuple(sorted(self._wires))

    @property
    def qubit_nb(self) -> int:
        return len(self._wires)

    @property
    def H(self) -> 'DAGCircuit':
        return DAGCircuit(Circuit(self).H)

    def run(self, ket: State) -> State:
        for elem in self:
            ket = elem.run(ket)
        return ket

    def evolve(self, rho):
        """Every operation applied to a Density in topological order
        (dagcircuit.py:86-89)."""
        for elem in self:
            rho = elem.evolve(rho)
        return rho

    def asgate(self) -> Gate:
        return Circuit(self).asgate()

    def aschannel(self):
        """The superoperator of the operations in order (dagcircuit.py:94-95)."""
        return Circuit(self).aschannel()

    def _depths(self, ops):
        """Layer index (0-based) of each op: 1 + max over its wire
        predecessors (dagcircuit.py:129-151)."""
        last: Dict[Qubit, int] = {}
        depth = []
        for elem in ops:
            d = max((last.get(q, -1) for q in elem.qubits), default=-1) + 1
            for q in elem.qubits:
                last[q] = d
            depth.append(d)
        return depth

    def depth(self, local: bool = True) -> int:
        """Longest path in operations (dagcircuit.py:97-113); with
        local=False only multi-qubit operations count."""
        ops = self._ops if local else [e for e in self._ops if len(e.qubits) > 1]
        d = self._depths(ops)
        return (max(d) + 1) if d else 0

    def size(self) -> int:
        return len(self._ops)

    def _component_sets(self):
        parent: Dict[Qubit, Qubit] = {q: q for q in self._wires}

        def find(q):
            while parent[q] != q:
                parent[q] = parent[parent[q]]
                q = parent[q]
            return q
        for elem in self._ops:
            qs = list(elem.qubits)
            for q in qs[1:]:
                a, b = find(qs[0]), find(q)
                if a != b:
                    parent[b] = a
        groups: Dict[Qubit, List[Qubit]] = {}
        for q in self._wires:
            groups.setdefault(find(q), []).append(q)
        return list(groups.values())

    def component_nb(self) -> int:
        """Number of independent sub-circuits (dagcircuit.py:119-122)."""
        return len(self._component_sets())

    def components(self) -> List['DAGCircuit']:
        """Split into independent sub-circuits (dagcircuit.py:124-127)."""
        comps = []
        for qs in self._component_sets():
            qset = set(qs)
            comps.append(DAGCircuit([e for e in self._ops if set(e.qubits) <= qset]))
        return comps

    def layers(self) -> Circuit:
        """A Circuit of Circuits, one per layer of mutually commuting
        operations on distinct qubits (dagcircuit.py:129-151)."""
        depth = self._depths(self._ops)
        nl = (max(depth) + 1) if depth else 0
        layers = [[] for _ in range(nl)]
        for elem, d in zip(self._ops, depth):
            layers[d].append(elem)
        return Circuit([Circuit(layer) for layer in layers])

    def __iter__(self) -> Iterator[Operation]:
        return iter(list(self._ops))

    def next_element(self, elem: Operation, qubit: Qubit = None):
        """Successor of elem on `qubit` (any wire if None) (dagcircuit.py:159-164)."""
        nxt = self._next[id(elem)]
        if qubit is None:
            return next(iter(nxt.values()))
        return nxt[qubit]

    def prev_element(self, elem: Operation, qubit: Qubit = None):
        """Predecessor of elem on `qubit` (dagcircuit.py:166-170)."""
        prv = self._prev[id(elem)]
        if qubit is None:
            return next(iter(prv.values()))
        return prv[qubit]
