"""Host-side logic that runs without a GPU: parameter mapping onto the C
structs, the generator's ground truth, and the oracle's stack loop on a tiny
stack (the CPU path the device is judged against)."""
import numpy as np

import lapstack_amd as la
from lapstack_amd import synth
from oracle import ref


def test_params_map_onto_the_c_layout():
    fp = la.FuseParams(levels=5, energy_radius=2, top_rule="average", use_chroma=True, depth_level=1, halo=2.0)._c()
    assert (fp.levels, fp.energy_radius, fp.top_rule, fp.use_chroma, fp.depth_level, fp.halo) == (5, 2, 2, 1, 1, 2.0)
    dp = la.DepthParams(scale=2, focus="sml", upsample="bilinear", tol=1e-6)._c()
    assert (dp.scale, dp.focus_kind, dp.upsample_kind) == (2, 1, 0)
    assert abs(dp.tol - 1e-6) < 1e-12
    assert la.DepthParams().grid(1024, 768) == (512, 384)
    assert la.DepthParams(scale=1).grid(8256, 5504) == (4128, 2752)
    assert la.AlignParams(model="affine").free() == [True, True, True, True, True, True, False, False]
    assert la.Interp.parse("spline36") == 4 and la.Interp.parse("lanczos") == 5


def test_generator_sims_are_the_jitter():
    b = synth.Bracket(96, 64, 4, seed=9, jitter=True)
    assert b.sims[0] == tuple(la.SIM_ID)
    for i in range(1, 4):
        s = b.sims[i]
        assert abs(s[0] * 96) <= 3 and abs(s[1] * 64) <= 3
        assert abs(s[2] - (1 + 0.002 * i)) < 1e-12 and abs(np.degrees(s[3])) <= 0.05
    f = b.frame(1).numpy()
    assert f.shape == (64, 96, 3) and f.dtype == np.uint16


def test_oracle_stack_recovers_generator_jitter():
    w, h, n = 256, 192, 3
    b = synth.Bracket(w, h, n, seed=4, jitter=True)
    frames = np.stack([b.frame(i).numpy() for i in range(n)])
    out = ref.stack_run(frames, ref.FuseParams.make(levels=4),
                        ref.StackParams.make(align=True, coarsen=0), ref.DepthParams.make())
    for i in range(1, n):
        m_got = ref.sim_matrix(out["sims"][i], w, h)
        m_true = ref.sim_matrix(np.array(b.sims[i]), w, h)
        for (x, y) in [(0, 0), (w - 1, h - 1), (0, h - 1), (w - 1, 0)]:
            a = ref.map_point(m_got, x, y)
            t = ref.map_point(m_true, x, y)
            assert abs(a[0] - t[0]) < 0.35 and abs(a[1] - t[1]) < 0.35, (i, out["sims"][i], b.sims[i])
    assert out["depth"].shape == (h, w) and np.isfinite(out["depth"]).all()
    assert 0 <= out["depth"].min() and out["depth"].max() <= n - 1
