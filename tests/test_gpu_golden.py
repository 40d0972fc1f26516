"""GPU parity with the reference's golden vectors (HIP path through the C ABI).

Each golden case (tests/golden/run_*.npz, produced by the compiled reference)
is replayed through libquillk.so:
  fp64 + global deposit   bit-level arithmetic of the reference up to the
                          order of the atomic current sums: 1e-10 relative
  fp32 + both deposits    per-component relative L2 field error <= 1e-4 and
                          particle moments within FP32 tolerance
"""
import numpy as np
import pytest

import golden_io as gio
import quill_amd as qa

pytestmark = pytest.mark.gpu

RUN_CASES = ["run_fdtd_boris", "run_ndfx_vay", "run_five_vay", "run_fdtd_vay_cold"]
SOLVER = {0: "fdtd", 1: "ndfx", 2: "five_point"}
PUSHER = {0: "vay", 1: "boris"}


def hdr(fx):
    return np.frombuffer(fx["header"].tobytes(), gio.HDR)[0]


def run_case(fx, precision, deposit):
    h = hdr(fx)
    print(f"case nx={int(h['nx'])} solver={SOLVER[int(h['solver'])]} pusher={PUSHER[int(h['pusher'])]} "
          f"precision={precision} deposit={deposit}", flush=True)
    r = qa.Region(int(h["nx"]), int(h["ny"]), int(h["nz"]), float(h["dx"]), float(h["dy"]), float(h["dz"]),
                  float(h["dt"]), solver=SOLVER[int(h["solver"])], pusher=PUSHER[int(h["pusher"])],
                  precision=precision, deposit=deposit, track_ids=True, e_s=float(h["e_s"]),
                  external_bz=float(h["external_bz"]), ion_cmr=[float(h["ion_cmr"][0])],
                  norm_n=float(h["norm_n"]), norm_ef=float(h["norm_ef"]), norm_ep=float(h["norm_ep"]))
    r.set_fields(fx["E_in"], fx["B_in"])
    r.add_records(fx["part_in"], fx["ids_in"])
    r.init_boundaries()
    diag = r.step(int(h["nsteps"]), diag=True)
    E, B, J, Be = r.get_fields()
    recs, ids = [], []
    for s in range(qa.ION0 + 1):   # electrons, positrons, photons, one ion population
        rec, idv = r.particles(s)
        recs.append(rec)
        ids.append(idv)
    r.close()
    rec = np.concatenate(recs)
    idv = np.concatenate(ids)
    order = np.argsort(idv, kind="stable")
    return dict(E=E, B=B, J=J, Be=Be, part=rec[order], ids=idv[order], diag=diag)


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / den if den > 0 else np.linalg.norm(a)


@pytest.mark.parametrize("name", RUN_CASES + ["run_edges"])
def test_fp64_global_matches_reference(name):
    fx = gio.load(name)
    out = run_case(fx, "fp64", "global")
    for key in ("E", "B", "J", "Be"):
        scale = np.abs(fx[key]).max()
        err = np.abs(out[key] - fx[key]).max() / scale
        assert err < 1e-10, f"{name} {key}: max rel err {err:.3e}"
    np.testing.assert_array_equal(out["ids"], fx["ids"])
    np.testing.assert_allclose(out["part"][:, :3], fx["part"][:, :3], rtol=0, atol=1e-10)
    np.testing.assert_allclose(out["part"][:, 3:6], fx["part"][:, 3:6], rtol=1e-9, atol=1e-12)
    ref = fx["diag"]
    for l, d in enumerate(out["diag"]):
        mine = [d[k] for k in gio.DIAG_NAMES]
        np.testing.assert_allclose(mine, ref[l], rtol=1e-10, atol=0, err_msg=f"{name} step {l}")


@pytest.mark.parametrize("deposit", ["global", "tiled"])
@pytest.mark.parametrize("name", RUN_CASES)
def test_fp32_matches_reference(name, deposit):
    fx = gio.load(name)
    out = run_case(fx, "fp32", deposit)
    for key in ("E", "B", "J", "Be"):
        for c in range(3):
            err = rel_l2(out[key][..., c], fx[key][..., c])
            assert err < 1e-4, f"{name} {key}[{c}]: rel L2 {err:.3e}"
    # particles: same survivors except those within rounding of a deletion face
    common, ia, ib = np.intersect1d(out["ids"], fx["ids"], return_indices=True)
    assert len(common) >= 0.995 * len(fx["ids"])
    a, b = out["part"][ia], fx["part"][ib]
    np.testing.assert_allclose(a[:, :3], b[:, :3], rtol=0, atol=2e-4)
    # momenta: per-particle vector error relative to |u| (FP32 rounding grows
    # through ten chaotic pushes in fields of amplitude 0.5)
    du = np.linalg.norm(a[:, 3:6] - b[:, 3:6], axis=1) / np.maximum(np.linalg.norm(b[:, 3:6], axis=1), 1e-3)
    assert np.median(du) < 1e-5 and np.quantile(du, 0.99) < 1e-3, (np.median(du), np.quantile(du, 0.99))
    # particle moments: charge-weighted momentum and kinetic energy
    for col in (3, 4, 5):
        m_a, m_b = (a[:, 7] * a[:, col]).sum(), (b[:, 7] * b[:, col]).sum()
        assert abs(m_a - m_b) <= 1e-4 * np.abs(b[:, 7] * b[:, col]).sum()
    ke_a, ke_b = (a[:, 7] * (a[:, 6] - 1)).sum(), (b[:, 7] * (b[:, 6] - 1)).sum()
    assert ke_a == pytest.approx(ke_b, rel=1e-5)
    ref = fx["diag"]
    for l, d in enumerate(out["diag"]):
        assert d["energy_f"] == pytest.approx(ref[l][6], rel=1e-5)
        assert d["n_e"] == pytest.approx(ref[l][3], rel=5e-3, abs=2)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_deposit_edge_cases_tiled(precision):
    """The hand-placed edge cases of run_edges.npz (ties between crossing
    fractions, starts on a face, ends within rounding of one, exits through
    every side, three crossings, eight particles in one cell; edge_case.py)
    through the tiled kernel: FP64 to 1e-10 of the reference, FP32 per field
    component within 1e-4 relative L2 and positions within 2e-4 cells."""
    fx = gio.load("run_edges")
    out = run_case(fx, precision, "tiled")
    if precision == "fp64":
        for key in ("E", "B", "J", "Be"):
            err = np.abs(out[key] - fx[key]).max() / np.abs(fx[key]).max()
            assert err < 1e-10, f"{key}: max rel err {err:.3e}"
        np.testing.assert_array_equal(out["ids"], fx["ids"])
        np.testing.assert_allclose(out["part"][:, :3], fx["part"][:, :3], rtol=0, atol=1e-10)
        np.testing.assert_allclose(out["part"][:, 3:6], fx["part"][:, 3:6], rtol=1e-9, atol=1e-12)
        return
    for key in ("E", "B", "J", "Be"):
        for c in range(3):
            err = rel_l2(out[key][..., c], fx[key][..., c])
            assert err < 1e-4, f"{key}[{c}]: rel L2 {err:.3e}"
    np.testing.assert_array_equal(out["ids"], fx["ids"])
    np.testing.assert_allclose(out["part"][:, :3], fx["part"][:, :3], rtol=0, atol=2e-4)
