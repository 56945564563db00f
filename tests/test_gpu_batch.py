"""Many registrations at once (clr_solve_gn_batch): every solve of the batch,
spread over several host threads and HIP streams, with and without captured
matvec graphs, equals the same solve run alone on the null stream, bit for
bit (all kernels and reductions are deterministic)."""
import numpy as np
import pytest

from claire_amd import Claire, solve_gn_batch, synth

pytestmark = pytest.mark.gpu

NT = 4
BETA = (1e-2, 1e-4, 0.0)
SIZES = [32, 32, 48, 32, 64, 48]


def _pair(n, seed):
    shape = (n, n, n)
    mT = synth.ellipsoid_phantom(shape, count=12, seed=seed, radii=(n / 16, n / 5))
    v = synth.bandlimited_velocity(shape, kmax=4, cfl=1.5, nt=NT, seed=seed + 100)
    return mT, v


def _problems(pairs, graphs):
    out = []
    for mT, v in pairs:
        c = Claire(mT.shape, nt=NT, regnorm="h1", beta=BETA, graphs=graphs)
        c.set_velocity(v)
        c.set_template_image(mT)
        mR = c.solve_state()[NT]
        c.set_reference_image(mR)
        out.append(c)
    return out


@pytest.mark.parametrize("graphs,workers", [(False, 3), (True, 4)])
def test_batch_equals_sequential(graphs, workers):
    print(f"\n[gpu case] batch graphs={graphs} workers={workers} sizes={SIZES}", flush=True)
    pairs = [_pair(n, 10 + i) for i, n in enumerate(SIZES)]
    seq = _problems(pairs, graphs=False)
    bat = _problems(pairs, graphs=graphs)
    try:
        want = [c.solve_gn(maxit=6) for c in seq]
        got = solve_gn_batch(bat, workers=workers, maxit=6)
        for i, ((v, rep), (vw, repw, _)) in enumerate(zip(got, want)):
            assert rep == repw, (i, rep, repw)
            assert np.array_equal(v, vw), i
            assert rep["iterations"] >= 1 and rep["J"] < rep["J0"]
        # contexts are back on their own streams and usable
        assert np.array_equal(bat[0].solve_gn(maxit=6)[0], want[0][0])
    finally:
        for c in seq + bat:
            c.close()
