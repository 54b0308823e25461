"""GPU: the measurement kernels of the stacked-load-case element contraction
(libfemgpu_probe.so, bench.py --workload contraction).

* The output layout of v_mfma_f32_4x4x1f32 that the matrix-instruction
  variant relies on: with A = lane and B = 100 + lane, lane 4b + j holds
  D[block b][row i][column j] = A(lane 4b + i) * B(lane 4b + j) in its
  register i (exact in FP32 for these small integers).
* Both variants give the same forces, and the same element energies
  u_case' F_case (the output the timing uses), on plates of the batch
  workload: the on-the-fly variant is the solver's tri_force_free (pinned
  by the matvec parity tests), the other forms K_e and contracts it with
  the matrix instruction, so the two differ by FP32 rounding only (bound
  5e-5 of the element's largest force)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mfma_4x4x1_layout():
    import torch
    from cudafemproject_amd import contraction_probe as CP
    out = torch.zeros(256, device="cuda", dtype=torch.float32)
    CP.mfma_layout(out)
    torch.cuda.synchronize()
    d = out.cpu().numpy().reshape(64, 4)
    lane = np.arange(64)
    blk, j = lane // 4, lane % 4
    want = np.stack([(4 * blk + i) * (100 + 4 * blk + j) for i in range(4)], axis=1)
    assert np.array_equal(d, want.astype(np.float32))


@pytest.mark.parametrize("n_plates", [3, 16])
def test_contraction_variants_agree(n_plates):
    import torch
    from cudafemproject_amd import contraction_probe as CP
    from cudafemproject_amd.problem import plate_batch
    insts = plate_batch(n_plates, seed=7)
    coords, elems = CP.flat_plates(insts)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((len(coords), CP.N_CASES, 2)).astype(np.float32)
    E, nu = 2e7, 0.3
    lam, mu = E * nu / (1 - nu * nu), E / (2 * (1 + nu))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    te, tc, tu = dev(elems), dev(coords), dev(u)
    m = len(elems)
    f0 = torch.full((m, CP.N_CASES, 6), float("nan"), device="cuda")
    f1 = torch.full((m, CP.N_CASES, 6), float("nan"), device="cuda")
    print(f"[contraction] {n_plates} plates, {m} triangles", flush=True)
    CP.contract(0, te, tc, tu, lam, mu, f0)
    CP.contract(1, te, tc, tu, lam, mu, f1)
    torch.cuda.synchronize()
    a, b = f0.cpu().numpy(), f1.cpu().numpy()
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    # per element and load case, relative to the largest force of the element
    scale = np.abs(a).max(axis=2, keepdims=True)
    err = float((np.abs(a - b) / scale).max())
    print(f"[contraction] max relative difference {err:.2e}", flush=True)
    assert err < 5e-5
    # each element's forces sum to zero (translation invariance of K_e)
    assert float((np.abs(b.reshape(m, CP.N_CASES, 3, 2).sum(axis=2)) / scale).max()) < 5e-5
    # energy mode: u_case' F_case per element and load case, both variants
    g0 = torch.full((m, CP.N_CASES), float("nan"), device="cuda")
    g1 = torch.full((m, CP.N_CASES), float("nan"), device="cuda")
    CP.contract(0, te, tc, tu, lam, mu, g0)
    CP.contract(1, te, tc, tu, lam, mu, g1)
    torch.cuda.synchronize()
    ga, gb = g0.cpu().numpy(), g1.cpu().numpy()
    ue = u[elems].transpose(0, 2, 1, 3).reshape(m, CP.N_CASES, 6)
    en = (ue.astype(np.float64) * a).sum(axis=2)
    # a force error bounded relative to the element's largest force bounds
    # the energy error relative to that force times sum |u|
    mag = scale[..., 0] * np.abs(ue.astype(np.float64)).sum(axis=2)
    assert np.all(np.isfinite(ga)) and np.all(np.isfinite(gb))
    assert float((np.abs(ga - en) / mag).max()) < 5e-5
    assert float((np.abs(gb - en) / mag).max()) < 5e-5
