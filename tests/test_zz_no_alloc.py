"""nb_advance_noalloc: the allocation-free variant of nb_advance, whose force
pass keeps its split partials in the library's static device scratch.

Bars:
  * bit-identical to nb_advance (same kernels, same split counts) for the
    mixed and FP64-fast precisions, on two calls in a row with the second on
    a smaller system (the scratch keeps the first call's contents, which the
    second must not depend on), and again after the context's own nb_advance;
  * partials larger than the scratch are refused with NB_ERR_CAPACITY, and a
    context without a direct KDK evolve with NB_ERR_NO_EVOLUTION, before any
    launch.
"""
import numpy as np
import pytest

from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu


def make(m, pos, vel, prec, steps=3):
    s = capi.System(0)
    s.set_G(1.0)
    s.set_timestep(1e-3)
    s.set_softening(0.01)
    s.set_precision(prec)
    s.set_iterations(steps)
    s.set_bodies(m, pos, vel)
    s.set_tuning("step_graphs", 2)
    return s


@pytest.mark.parametrize("prec", [capi.PREC_MIXED, capi.PREC_FP64_FAST])
def test_noalloc_matches_advance(prec):
    for n in (20000, 3000):  # second call smaller than the first
        print(f"case noalloc prec={prec} n={n}", flush=True)
        m, pos, vel = capi.ic_plummer(60 + n % 7, n)
        a = make(m, pos, vel, prec)
        b = make(m, pos, vel, prec)
        a.evolve(capi.LEAPFROG_KDK)
        b.evolve(capi.LEAPFROG_KDK)
        a.advance(9)
        b.advance_noalloc(9)
        sa, sb = a.state(), b.state()
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(sa[k], sb[k]), (n, k)
        assert a.time_elapsed() == b.time_elapsed()
        # the two entry points interleave on one context
        a.advance(4)
        b.advance(2)
        b.advance_noalloc(2)
        sa, sb = a.state(), b.state()
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(sa[k], sb[k]), (n, k)
        a.close()
        b.close()


def test_noalloc_errors():
    m, pos, vel = capi.ic_plummer(61, 65536)
    s = make(m, pos, vel, capi.PREC_MIXED, steps=1)
    with pytest.raises(capi.NBodyError) as e:
        s.advance_noalloc(1)  # nothing evolved yet
    assert e.value.code == capi.NB_ERR_NO_EVOLUTION
    # 4 i-bodies per lane at 16 waves per SIMD splits j 64 ways:
    # 64 x 6 x 65536 floats = 96 MiB of partials, beyond the 32 MiB scratch
    s.set_tuning("direct_ipl", 4)
    s.set_tuning("direct_waves", 16)
    s.evolve(capi.LEAPFROG_KDK)
    before = s.state()["pos"]
    with pytest.raises(capi.NBodyError) as e:
        s.advance_noalloc(1)
    assert e.value.code == capi.NB_ERR_CAPACITY
    assert np.array_equal(s.state()["pos"], before)  # refused before any launch
    s.set_precision(capi.PREC_FP64)
    s.evolve(capi.LEAPFROG_KDK)
    with pytest.raises(capi.NBodyError) as e:
        s.advance_noalloc(1)
    assert e.value.code == capi.NB_ERR_NOT_SUPPORTED
    s.close()
