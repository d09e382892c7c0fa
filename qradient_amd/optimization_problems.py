"""Problem adapters (reference qradient/optimization_problems.py:3-37)."""
import numpy as np


class MaxCut:
    """MaxCut(vertex_num, edge_set=...) -> observable dict (optimization_problems.py:3-37).

    Random edge sets (edge_num / edge_probability) draw from the global numpy
    RNG in the reference's order; ``regular3(n, rng)`` adds seeded random
    3-regular graphs, the input family of the batched QAOA workload.
    """

    # edge sources by precedence: an explicit set wins, then a random set of
    # a fixed size, then a random set with a per-edge probability
    _SOURCES = ('edge_set', 'edge_num', 'edge_probability')

    def __init__(self, vertex_num, **kwargs):
        self.vertex_num = vertex_num
        source = next((k for k in self._SOURCES if k in kwargs), None)
        if source is None:
            raise ValueError('Specify one of the three edge_set, edge_num, or edge_probability')
        self.edge_set = kwargs[source] if source == 'edge_set' else self._draw(source, kwargs[source])

    def _draw(self, source, value):
        """Random edges from the global numpy RNG.  Candidate pairs (i, j),
        i < j, in row-major order; edge_num: one choice without replacement
        over their indices; edge_probability: one uniform draw per pair, in
        that order (the reference's draw sequence, so seeded graphs agree)."""
        pairs = np.stack(np.triu_indices(self.vertex_num, 1), axis=1)
        if source == 'edge_num':
            return pairs[np.random.choice(range(len(pairs)), size=value, replace=False)]
        keep = np.array([np.random.rand() < value for _ in range(len(pairs))], dtype=bool)
        return pairs[keep]

    def to_observable(self):
        """{'zz': n x n object array with 1.0 at [a, b]} (optimization_problems.py:33-37)."""
        zz = np.full((self.vertex_num, self.vertex_num), None)
        edges = np.asarray(self.edge_set, dtype=np.int64).reshape(-1, 2)
        zz[edges[:, 0], edges[:, 1]] = 1.
        return {'zz': zz}

    @staticmethod
    def regular3(n, rng):
        """Seeded pairing-model random 3-regular simple graph, edges (a, b) with a < b."""
        if n % 2:
            raise ValueError('a 3-regular graph needs an even vertex number')
        while True:
            stubs = np.repeat(np.arange(n), 3)
            rng.shuffle(stubs)
            pairs = stubs.reshape(-1, 2)
            seen = set()
            ok = True
            for a, b in pairs:
                e = (min(a, b), max(a, b))
                if a == b or e in seen:
                    ok = False
                    break
                seen.add(e)
            if ok:
                return MaxCut(n, edge_set=np.array(sorted(seen)))
