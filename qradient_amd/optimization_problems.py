"""Problem adapters (reference qradient/optimization_problems.py:3-37)."""
import numpy as np


class MaxCut:
    """MaxCut(vertex_num, edge_set=...) -> observable dict (optimization_problems.py:3-37).

    Random edge sets (edge_num / edge_probability) draw from the global numpy
    RNG in the reference's order; ``regular3(n, rng)`` adds seeded random
    3-regular graphs, the input family of the batched QAOA workload.
    """

    def __init__(self, vertex_num, **kwargs):
        self.vertex_num = vertex_num
        if 'edge_set' in kwargs:
            self.edge_set = kwargs['edge_set']
        elif 'edge_num' in kwargs:
            self.__random(edge_num=kwargs['edge_num'])
        elif 'edge_probability' in kwargs:
            self.__random(edge_probability=kwargs['edge_probability'])
        else:
            raise ValueError('Specify one of the three edge_set, edge_num, or edge_probability')

    def __random(self, **kwargs):
        all_edges = np.array([[i, j] for i in range(self.vertex_num) for j in range(i + 1, self.vertex_num)])
        if 'edge_num' in kwargs:
            idx = np.random.choice(range(len(all_edges)), size=kwargs['edge_num'], replace=False)
            self.edge_set = all_edges[idx]
        else:
            self.edge_set = np.array([e for e in all_edges if np.random.rand() < kwargs['edge_probability']])

    def to_observable(self):
        """{'zz': n x n object array with 1.0 at [a, b]} (optimization_problems.py:33-37)."""
        zz = np.full([self.vertex_num, self.vertex_num], None)
        for edge in self.edge_set:
            zz[edge[0], edge[1]] = 1.
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
