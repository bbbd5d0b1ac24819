"""Analytic regularization on the GPU (plm_hip_bspline_regularize) against
the oracle's restatement of Bspline_regularize flavor c
(src/register/bspline_regularize_analytic.cxx:638-1007).

Both are double precision.  The GPU applies each Kronecker operator
separably (4x4 contractions along the knot digits) and groups equal
monomials, the oracle multiplies by the dense 64x64 matrices the
reference builds: the same sums in another order, so they agree to ~1e-13
relative.  The oracle itself is pinned by closed forms (constant and
linear fields), finite differences and a dense numpy evaluation
(tests/test_oracle.py)."""
import numpy as np
import pytest

import oracle
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume
from tests.helpers import small_pair

pytestmark = pytest.mark.gpu

WEIGHTS = [
    dict(curvature_penalty=1.1),
    dict(diffusion_penalty=0.7, total_displacement_penalty=0.3),
    dict(third_order_penalty=0.2, curvature_penalty=0.5, curvature_mixed_weight=0.6),
    dict(lame_coefficient_1=0.4, lame_coefficient_2=0.9, linear_elastic_multiplier=1.3),
    dict(total_displacement_penalty=0.3, diffusion_penalty=0.7, curvature_penalty=1.1,
         lame_coefficient_1=0.4, lame_coefficient_2=0.9, linear_elastic_multiplier=1.3,
         third_order_penalty=0.2, curvature_mixed_weight=0.6),
]


def stage(dim=(40, 36, 28), spacing=(1.0, 1.25, 2.5), vpr=(5, 6, 4), roi=None):
    pair = small_pair(dim=dim, spacing=spacing, vpr=vpr, seed=3)
    fixed = Volume(pair.fixed, pair.spacing)
    moving = Volume(pair.moving, pair.spacing)
    kw = {"roi_offset": roi[0], "roi_dim": roi[1]} if roi else {}
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn, **kw)
    st = BsplineHip(bxf, fixed, moving)
    g = oracle.geom(pair.dim, pair.spacing, pair.vox_per_rgn, **kw)
    return bxf, st, g


@pytest.mark.parametrize("wi", range(len(WEIGHTS)))
@pytest.mark.parametrize("roi", [None, ((3, 4, 2), (31, 27, 22))])
def test_regularize_matches_oracle(hip_lib, wi, roi):
    import torch
    bxf, st, g = stage(roi=roi)
    w = WEIGHTS[wi]
    rng = np.random.default_rng(10 + wi)
    coeff = rng.normal(0, 2.0, bxf.num_coeff).astype(np.float32)
    d_c = torch.from_numpy(coeff).cuda()
    sc = torch.zeros(1, dtype=torch.float64, device="cuda")
    gr = torch.full((bxf.num_coeff,), 5.0, dtype=torch.float64, device="cuda")
    st.regularize(d_c, sc, gr, **w)
    torch.cuda.synchronize()
    s_ref, g_ref = oracle.reg_analytic(g, oracle.reg_parms(**w), coeff)
    s_got, g_got = sc.item(), gr.cpu().numpy()
    print(f"weights {w}: S {s_got:.6e} vs {s_ref:.6e}")
    assert abs(s_got - s_ref) <= 1e-11 * abs(s_ref)
    assert np.abs(g_got - g_ref).max() <= 1e-11 * np.abs(g_ref).max()
    # accumulate adds to what the gradient holds
    st.regularize(d_c, sc, gr, accumulate=True, **w)
    torch.cuda.synchronize()
    assert np.abs(gr.cpu().numpy() - 2 * g_ref).max() <= 1e-11 * np.abs(g_ref).max()
    st.close()


def test_regularize_zero_weights_and_deterministic(hip_lib):
    import torch
    bxf, st, g = stage()
    coeff = torch.from_numpy(np.random.default_rng(1).normal(0, 2, bxf.num_coeff).astype(np.float32)).cuda()
    sc = torch.zeros(1, dtype=torch.float64, device="cuda")
    gr = torch.zeros(bxf.num_coeff, dtype=torch.float64, device="cuda")
    st.regularize(coeff, sc, gr)
    torch.cuda.synchronize()
    assert sc.item() == 0.0 and not torch.any(gr != 0)
    outs = []
    for _ in range(2):
        st.regularize(coeff, sc, gr, **WEIGHTS[-1])
        torch.cuda.synchronize()
        outs.append((sc.item(), gr.cpu().numpy().copy()))
    assert outs[0][0] == outs[1][0] and np.array_equal(outs[0][1], outs[1][1])
    st.close()
