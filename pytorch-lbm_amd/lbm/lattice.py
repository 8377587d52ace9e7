"""D2Q9 / D3Q19 lattices (lbm/lattice.py:10-96 of the reference).

Same direction numbering, fp32 weights (kept fp32 even for fp64 runs, as the
reference does at lbm/lattice.py:21) and opposite tables; the device kernels
hold identical tables in pytorch-lbm_amd/csrc/lbm_lattice.hpp.
"""
import numpy as np
import torch


class Lattice:
    def __init__(self, e, w, opp, cs2=1.0 / 3.0):
        self.dim = e.shape[1]
        self.q = e.shape[0]
        self.cs2 = cs2
        self.e = torch.tensor(e, dtype=torch.float32)
        self.w = torch.tensor(w, dtype=torch.float32)
        self.opp = torch.tensor(opp, dtype=torch.long)
        self.ew = self.e * self.w.unsqueeze(1)
        self.e_sq = (self.e ** 2).sum(dim=1)
        self.stream_shifts = -np.asarray(e, dtype=int)


def _d2q9():
    e = np.array([(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1),
                  (1, 1), (-1, 1), (-1, -1), (1, -1)], dtype=np.int32)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float32)
    opp = np.array([0, 2, 1, 4, 3, 7, 8, 5, 6], dtype=np.int64)
    return e, w, opp


def _d3q19():
    faces = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    edges = [(1, 1, 0), (-1, 1, 0), (-1, -1, 0), (1, -1, 0),
             (1, 0, 1), (-1, 0, 1), (-1, 0, -1), (1, 0, -1),
             (0, 1, 1), (0, -1, 1), (0, -1, -1), (0, 1, -1)]
    e = np.array([(0, 0, 0)] + faces + edges, dtype=np.int32)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
    opp = np.array([int(np.flatnonzero((e == -v).all(1))[0]) for v in e], dtype=np.int64)
    return e, w, opp


class D2Q9(Lattice):
    def __init__(self):
        super().__init__(*_d2q9())


class D3Q19(Lattice):
    def __init__(self):
        super().__init__(*_d3q19())
