"""Host logic of the DNS-aided quantities (no GPU): the filter pieces the
host mirror hands to ec_filter_coarse for apply_DHpi, get_sigDH and get_rDHi
(src/N.jl:1038-1115, 1228-1237), built from get_kernels' centred kernels and
the coarse-grain positions, equal the oracle's independently composed
pieces (face / box / centre / point pieces convolved with the Gaussian)."""
import numpy as np
import pytest


@pytest.mark.parametrize("c,dr", [(3, 0.0), (5, 2.0), (7, 1.5), (3, 2.0)])
def test_dnsaid_quantities_match_oracle_pieces(ec, O, c, dr):
    NS = ec.navierstokes
    n = 12 * c
    g, gl = NS.Grid(1.0, n), NS.Grid(1.0, n // c)
    FD = NS.gaussian(g, dr * NS.spacing(gl), 2)
    FDH = NS.composekernel(NS.tophat(g, c), FD)
    mine = {name: (src, pcs) for name, src, pcs in NS.dnsaid_quantities(c, FD, FDH)}
    ref = O.closure_filters(c, "va", O.gaussian(1.0 / n, dr / (n // c), 2))
    assert set(mine) == set(ref)
    for name, (src, pcs) in ref.items():
        msrc, mpcs = mine[name]
        assert msrc == src, name
        for (w, off), (mw, moff) in zip(pcs, mpcs):
            assert off == moff and len(w) == len(mw), name
            assert np.abs(np.asarray(w) - np.asarray(mw)).max() < 1e-15, name


def test_stress_positions(ec):
    """lazytensorcoarsegrain positions: centre for 11, 22, 33; edges (T,T,F),
    (F,T,T), (T,F,T) for 12, 23, 31 (one-tap kernels at c J - !stag c/2)."""
    NS = ec.navierstokes
    c = 5
    got = {name: [off for _, off in pcs] for name, _, pcs in NS.stress_positions(c)}
    a = -(c // 2)
    assert got == {"s11": [a, a, a], "s22": [a, a, a], "s33": [a, a, a], "s12": [0, 0, a], "s23": [a, 0, 0],
                   "s31": [0, a, 0]}
