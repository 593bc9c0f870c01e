"""A synthetic no-fusion tensordot case: rank-3 blocks of A (a0, a1, a2) and B (b0, b1, b2), contracted
over A's legs (0, 2) with B's legs (1, 0); permutations Aorder = (1, 0, 2), Border = (1, 0, 2); every
output block sums two or three block products (test infrastructure)."""
import numpy as np


def make(rng, cplx_a=False, cplx_b=False):
    # contracted dims (c0, c2) per term and the free dims
    terms = [((2, 3), (4, 5)), ((3, 2), (4, 5)), ((1, 4), (4, 5)), ((2, 2), (6, 3)), ((5, 1), (6, 3)),
             ((3, 3), (2, 7)), ((2, 1), (2, 7)), ((4, 2), (2, 7))]
    groups = [[0, 1, 2], [3, 4], [5, 6, 7]]
    Areshape, Breshape, a_parts, b_parts = [], [], [], []
    ao = bo = 0
    for (c0, c2), (m, n) in terms:
        Di = (c0, m, c2)  # A block legs (0, 1, 2); Aorder (1, 0, 2) -> (m, c0, c2) -> (m, c0 c2)
        Areshape.append(((ao, ao + c0 * m * c2), Di, m, c0 * c2))
        a = rng.standard_normal(Di) + (1j * rng.standard_normal(Di) if cplx_a else 0)
        a_parts.append(a.ravel())
        ao += c0 * m * c2
        Dj = (c2, c0, n)  # B legs (0, 1, 2) = (c2, c0, n); Border (1, 0, 2) -> (c0, c2, n) -> (c0 c2, n)
        Breshape.append(((bo, bo + c2 * c0 * n), Dj, c0 * c2, n))
        b = rng.standard_normal(Dj) + (1j * rng.standard_normal(Dj) if cplx_b else 0)
        b_parts.append(b.ravel())
        bo += c2 * c0 * n
    meta_dot, co = [], 0
    for grp in groups:
        m, n = terms[grp[0]][1]
        meta_dot.append(((co, co + m * n), (m, n), tuple((t, t) for t in grp)))
        co += m * n
    return (np.concatenate(a_parts), np.concatenate(b_parts), tuple(meta_dot), tuple(Areshape), tuple(Breshape),
            (1, 0, 2), (1, 0, 2), co)
