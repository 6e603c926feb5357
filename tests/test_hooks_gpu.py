"""Step hooks (nb_set_step_hook) and in-place state updates (nb_update_state):
the engine side of evolve's for_all_bods / for_each_bod callbacks
(gregsys.hpp:2593-2607, 2946-2990).

Bars: hooks run once per step in the reference's order (NB_HOOK_STEP after
the step's update and before the collision pass, NB_HOOK_BODIES after it) for
every direct-sum integrator and never for Barnes-Hut; a state change made in
a hook equals the same change made between two evolve calls, bit for bit; a
failing hook stops the evolve with NB_ERR_CALLBACK; re-entrant stepping from
a hook is refused.
"""
import numpy as np
import pytest

from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
G_SI = 6.6743e-11


def make(m, pos, vel, steps, prec=capi.PREC_FP64, r=None, e=None, coll=False):
    s = capi.System(0)
    s.set_G(G_SI)
    s.set_timestep(10.0)
    s.set_softening(1e6)
    s.set_iterations(steps)
    s.set_precision(prec)
    if coll:
        s.set_collisions(capi.OVERLAP_COLL_CHECK)
    s.set_bodies(m, pos, vel, r=r, rest_c=e)
    return s


@pytest.mark.parametrize("method", [capi.LEAPFROG_KDK, capi.EULER, capi.MIDPOINT, capi.LEAPFROG_DKD])
def test_hook_order(method):
    m, pos, vel = capi.ic_uniform_ball(31, 64, 1e20, 1e22, 1e9, 1e3)
    s = make(m, pos, vel, 7)
    calls = []
    s.set_step_hook(lambda step, where: calls.append((step, where)) or 0)
    s.evolve(method)
    assert calls == [(k, w) for k in range(1, 8) for w in (capi.HOOK_STEP, capi.HOOK_BODIES)]
    calls.clear()
    s.advance(3)
    assert calls == [(k, w) for k in range(8, 11) for w in (capi.HOOK_STEP, capi.HOOK_BODIES)]
    s.set_step_hook(None)
    calls.clear()
    s.advance(2)
    assert calls == []


def test_no_hooks_for_barnes_hut():
    m, pos, vel = capi.ic_plummer(32, 2000)
    s = capi.System(0)
    s.set_G(1.0)
    s.set_timestep(1e-3)
    s.set_softening(0.01)
    s.set_iterations(3)
    s.set_bodies(m, pos, vel)
    calls = []
    s.set_step_hook(lambda step, where: calls.append(step) or 0)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
    assert calls == []


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED, capi.PREC_FP64_FAST])
def test_change_in_hook_equals_change_between_calls(prec):
    """Velocities shifted and restitution rewritten by a hook at step 5 of 10,
    against evolve(5), the same change through nb_update_state, advance(5)."""
    print(f"case hook change prec={prec}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(33, 300, 1e20, 1e22, 1e9, 1e3)
    dv = np.full((300, 3), 3.0)
    a = make(m, pos, vel, 10, prec)

    def hook(step, where):
        if step == 5 and where == capi.HOOK_STEP:
            st = a.state()
            a.update_state(vel=st["vel"] + dv, rest_c=np.full(300, 0.25))
        return 0

    a.set_step_hook(hook, capi.HOOK_STEP)
    a.evolve()
    b = make(m, pos, vel, 5, prec)
    b.evolve()
    st = b.state()
    b.update_state(vel=st["vel"] + dv, rest_c=np.full(300, 0.25))
    b.advance(5)
    sa, sb = a.state(), b.state()
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(sa[k], sb[k]), k


def test_hook_failure_and_reentry():
    m, pos, vel = capi.ic_uniform_ball(34, 50, 1e20, 1e22, 1e9, 1e3)
    s = make(m, pos, vel, 10)
    codes = []

    def hook(step, where):
        try:
            s.advance(1)
        except capi.NBodyError as e:
            codes.append(e.code)
        return 7 if step == 3 else 0

    s.set_step_hook(hook, capi.HOOK_BODIES)
    with pytest.raises(capi.NBodyError) as e:
        s.evolve()
    assert e.value.code == capi.NB_ERR_CALLBACK
    assert codes == [capi.NB_ERR_INVALID_ARGUMENT] * 3
    s.set_step_hook(None)
    s.evolve()  # the context stays usable
    with pytest.raises(capi.NBodyError) as e:
        s.update_state(m=-np.ones(50))
    assert e.value.code == capi.NB_ERR_NEGATIVE_MASS


def test_hooks_with_collisions_see_merges():
    """With collisions on, NB_HOOK_BODIES sees the body count after the pass."""
    rng = np.random.default_rng(1)
    n = 24
    m = 10 ** rng.uniform(11, 13, n)
    r = (3 * m / (4 * np.pi * 500)) ** (1 / 3)
    pos = rng.uniform(-6e3, 6e3, (n, 3))
    vel = rng.normal(0, 0.5, (n, 3))
    s = capi.System(0)
    s.set_G(G_SI)
    s.set_timestep(5.0)
    s.set_iterations(300)
    s.set_collisions(capi.OVERLAP_COLL_CHECK, mergers=True)
    s.set_bodies(m, pos, vel, r=r, rest_c=rng.uniform(0.1, 0.9, n))
    counts = []
    s.set_step_hook(lambda step, where: counts.append((where, s.num_bodies())) or 0)
    s.evolve()
    after = [c for w, c in counts if w == capi.HOOK_BODIES]
    before = [c for w, c in counts if w == capi.HOOK_STEP]
    assert len(after) == 300 and after[-1] == s.num_bodies() < n
    assert any(b > a for b, a in zip(before, after))  # some pass merged bodies
