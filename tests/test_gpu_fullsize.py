"""GPU: size-independent properties of the H_eff matvec, eigensolve and
adjoint at the benchmarked size (d = 64 Z2 coupled chain, chi = 500,
n = 5.12e8; the oracle cannot run there), on the autotuned launches the
benchmark uses.  Vectors cross the C ABI as host arrays (8 GB each).

* Hermiticity <x, H y> = <H x, y> (1e-12) and linearity (1e-13) of the FP64
  matvec on the benchmark's Hermitian environments; the FP32 split-complex
  matvec against FP64 (2e-5); the sliced matvec (2 x 2) against the
  unsliced one (1e-12);
* one a0' row of the output against an independent host contraction of the
  same row (numpy tensordot over L, W, W, R; 1e-12);
* environments of a real iDMRG run grown to chi = 500: the converged
  FP32-mixed lowest eigenvalue within 1e-10 of the converged FP64 one, with
  the state overlap reported; both eigenvectors are eigenvectors of the
  FP64 matvec;
* adjoint on that H_eff: dE/dpsi = (H psi - E psi)/|psi|^2 (1e-12) and, at
  the ground state, dE/dJ_perp and dE/dE_n (random directions) against
  exact central differences of <psi|H(J)|psi> built on the device.
"""
import numpy as np
import pytest

from conftest import from_lib, rel_err, to_lib

pytestmark = pytest.mark.gpu

CHI, D = 500, 64


@pytest.fixture()
def bench_heff():
    """The benchmark's H_eff, on a context of its own (its workspaces are
    released with it at the end of the test)."""
    import chainamps_amd as cam
    from chainamps_amd import workloads
    c = cam.Context(0)
    L, W, R = workloads.coupled_chain_heff_inputs(c, chi=CHI, d=D, seed=0, Jperp=0.1, hermitian=True)
    H = cam.Heff(L, W, R)
    yield c, H, (L, W, R)
    H.close()
    for t in (L, W, R):
        t.close()
    c.close()


def _vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_fullsize_hermitian_linear_fp32_sliced(bench_heff):
    print("case fullsize chi=500 properties", flush=True)
    import chainamps_amd as cam
    c, H, _ = bench_heff
    n = H.n
    assert n == (CHI * D) ** 2 // 2
    x, y = _vec(n, 1), _vec(n, 2)
    Hx = H.apply(x)                  # the first call autotunes
    Hy = H.apply(y)
    print(f"  launch fp64 {H.get_launch(cam.CAM_FP64)}", flush=True)
    a, b = np.vdot(x, Hy), np.vdot(Hx, y)
    scale = np.linalg.norm(Hx) * np.linalg.norm(y)
    print(f"  hermiticity |<x,Hy>-<Hx,y>|/(|Hx||y|) = {abs(a - b) / scale:.3e}", flush=True)
    assert abs(a - b) <= 1e-12 * scale
    al, be = 0.3 - 1.1j, 2.0 + 0.5j
    err = rel_err(H.apply(al * x + be * y), al * Hx + be * Hy)
    print(f"  linearity {err:.3e}", flush=True)
    assert err <= 1e-13
    del y, Hy
    err32 = rel_err(H.apply(x, cam.CAM_FP32), Hx)
    print(f"  fp32 vs fp64 {err32:.3e}; launch fp32 {H.get_launch(cam.CAM_FP32)}", flush=True)
    assert err32 <= 2e-5
    H.set_slices(2, 2)
    errs = rel_err(H.apply(x), Hx)
    print(f"  sliced 2x2 vs unsliced {errs:.3e}", flush=True)
    assert errs <= 1e-12


def test_fullsize_row_against_host_contraction(bench_heff):
    """Output entries of one a0' row from numpy tensordots over the dense
    L, W, R and psi (host), at the full size."""
    print("case fullsize chi=500 host row", flush=True)
    c, H, (L, W, R) = bench_heff
    n = H.n
    v = _vec(n, 5)
    out = H.apply(v)
    print("  device matvec done", flush=True)
    allowed = H.allowed()
    Ld = from_lib(L, None).to_dense_tensor()[0, :, 0, :, 0, :]     # (a0', b, a0)
    Rd = from_lib(R, None).to_dense_tensor()[:, 0, :, 0, :, 0]     # (a2', b, a2)
    Wd = from_lib(W, None).to_dense_tensor()                       # (s', bl, s, br)
    psi = np.zeros(D * CHI * D * CHI, np.complex128)
    psi[allowed] = v
    del v
    psi = psi.reshape((D, CHI, D, CHI), order="F")                 # (s1, a0, s2, a2)
    a0p = 137
    t1 = np.tensordot(Ld[a0p], psi, axes=([1], [1]))              # (b, s1, s2, a2)
    del psi
    t2 = np.tensordot(t1, Wd, axes=([0, 1], [1, 2]))              # (s2, a2, s1', b1)
    t3 = np.tensordot(t2, Wd, axes=([3, 0], [1, 2]))              # (a2, s1', s2', b2)
    row = np.tensordot(t3, Rd, axes=([0, 3], [2, 1]))             # (s1', s2', a2')
    print("  host row done", flush=True)
    # entries of the device output in that row: flat = s1 + D a0 + D CHI (s2 + D a2)
    sel = np.nonzero((allowed // D) % CHI == a0p)[0]
    f = allowed[sel]
    s1 = f % D
    rest = f // (D * CHI)
    s2, a2 = rest % D, rest // D
    got = out[sel]
    want = row[s1, s2, a2]
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"  row a0'={a0p}: {len(sel)} entries, relative error {err:.3e}", flush=True)
    assert len(sel) == D * D * CHI // 2
    assert err <= 1e-12


@pytest.fixture(scope="module")
def idmrg_heff():
    """H_eff of an iDMRG growth step at chi = 500 of the d=64 chain (the
    blocks of three FP64 growth steps), on a context of its own."""
    import chainamps_amd as cam
    from oracle import models
    c = cam.Context(0)
    rules, phys, W, E, sig = models.coupled_chain_model(d=D, seed=0, Jperp=0.1)
    Wt = to_lib(c, W)
    dm = cam.IDMRG(Wt, converge=False)
    for _ in range(3):
        r = dm.grow(CHI)
        print(f"  idmrg step N={r['N']} chi={r['chi']} {r['seconds']:.2f} s", flush=True)
    LB, RB = dm.blocks()
    dm.close()
    c.release_workspaces()
    H = cam.Heff(LB, Wt, RB)
    state = dict(ctx=c, H=H, LB=LB, RB=RB, W=Wt, E=E)
    yield state
    H.close()
    c.close()


def test_fullsize_mixed_vs_fp64_converged(idmrg_heff):
    print("case fullsize chi=500 converged mixed vs fp64", flush=True)
    H = idmrg_heff["H"]
    E64, x, i64 = H.lowest(converge=True, tol=1e-9)
    print(f"  fp64 E = {E64:.13f} {i64}", flush=True)
    E32, y, i32 = H.lowest(converge=True, tol=1e-9, mixed=True)
    print(f"  mixed E = {E32:.13f} {i32}", flush=True)
    # state error as the distance between the unit vectors after aligning
    # their phases (1 - |<x|y>| = dist^2 / 2 without the cancellation of
    # a 5e8-term dot product near 1)
    c = np.vdot(x, y)
    dist = np.linalg.norm(x / np.linalg.norm(x) - (c / abs(c)).conjugate() * y / np.linalg.norm(y))
    print(f"  |E_mixed - E_fp64| = {abs(E32 - E64):.3e}, state distance |x - e^(i phi) y| = {dist:.3e} "
          f"(1 - |<x|y>| = {dist * dist / 2:.3e})", flush=True)
    assert abs(E32 - E64) <= 1e-10 * max(1.0, abs(E64))
    assert dist <= 1.4e-4
    assert i32["matvecs_fp32"] > 0
    # both are eigenvectors of the FP64 matvec (the Krylov basis is freed
    # first: the host-vector matvec needs its own device buffers)
    idmrg_heff["ctx"].release_workspaces()
    for v, E in ((x, E64), (y, E32)):
        r = np.linalg.norm(H.apply(v) - E * v) / np.linalg.norm(v)
        assert r <= 1e-7 * abs(E), r
    idmrg_heff["psi0"] = x


def test_fullsize_adjoint(idmrg_heff):
    print("case fullsize chi=500 adjoint", flush=True)
    import chainamps_amd as cam
    c, H, LB, RB, Wt = (idmrg_heff[k] for k in ("ctx", "H", "LB", "RB", "W"))
    n = H.n
    rows, cols, p, i, x = Wt.to_csc()
    nnz = len(x)
    # dE/dpsi for a random psi (reference compressed order on the host)
    psi = _vec(n, 4)
    dpsi, dW = H.vjp(psi, nnz)
    Hpsi = H.apply(psi)
    n2 = np.vdot(psi, psi).real
    En = np.vdot(psi, Hpsi).real / n2
    err = rel_err(dpsi, (Hpsi - En * psi) / n2)
    print(f"  dE/dpsi relative error {err:.3e}", flush=True)
    assert err <= 1e-12
    del psi, dpsi, Hpsi
    # at the ground state: couplings against exact central differences of
    # <psi|H(W)|psi>/<psi|psi>, each H(W) built and applied on the device
    if "psi0" not in idmrg_heff:
        pytest.skip("needs the converged ground state of the previous test")
    psi0 = idmrg_heff["psi0"]
    dpsi, dW = H.vjp(psi0, nnz)
    print(f"  |dE/dpsi| at the ground state {np.linalg.norm(dpsi):.3e}", flush=True)
    col_of = np.repeat(np.arange(cols), np.diff(p))
    bl, br = i // D, col_of // D
    J = 0.1
    dEdJ = np.sum(dW * np.where((bl == 1) & (br == 0), x / J, 0))
    assert abs(dEdJ.imag) <= 1e-10 * max(1.0, abs(dEdJ))

    def energy(Wm):
        Hm = cam.Heff(LB, Wm, RB)
        out = Hm.apply(psi0)
        Hm.close()
        Wm.close()
        return (np.vdot(psi0, out) / np.vdot(psi0, psi0)).real

    def mpo(J=0.1, E=None):
        from oracle import models
        rules, phys, W, E0, sig = models.coupled_chain_model(d=D, seed=0, Jperp=J, E=E)
        return to_lib(c, W)
    h = 1e-3      # <psi|H|psi> is linear in J and in the level energies: exact differences
    fd = (energy(mpo(J + h)) - energy(mpo(J - h))) / (2 * h)
    print(f"  dE/dJ adjoint {dEdJ.real:.12f}, central difference {fd:.12f}", flush=True)
    assert abs(dEdJ.real - fd) <= 1e-8 * max(1.0, abs(fd))
    lev = np.zeros(D)
    for nlev in range(D):
        k = np.where((bl == 2) & (br == 0) & (i % D == nlev) & (col_of % D == nlev))[0]
        assert len(k) == 1
        lev[nlev] = dW[k[0]].real
    # the library's chain rule (cam_heff_coupling_grad) against the assembly above
    from chainamps_amd import workloads
    gens = workloads.coupled_chain_generators((rows, cols, p, i, x), D, seed=0)
    g = H.coupling_grad(psi0, nnz, gens)
    print(f"  coupling_grad dE/dJ {g[0].real:.12f}, max |Im| {np.max(np.abs(g.imag)):.2e}", flush=True)
    assert abs(g[0] - dEdJ) <= 1e-12 * max(1.0, abs(dEdJ))
    assert np.max(np.abs(g[1:] - lev)) <= 1e-12 * max(1.0, np.max(np.abs(lev)))
    assert np.max(np.abs(g.imag)) <= 1e-10 * max(1.0, np.max(np.abs(g)))
    E = idmrg_heff["E"]
    rng = np.random.default_rng(11)
    for trial in range(2):
        u = rng.standard_normal(D)
        ref = (energy(mpo(E=E + h * u)) - energy(mpo(E=E - h * u))) / (2 * h)
        print(f"  dE/dE_n . u adjoint {lev @ u:.12f}, central difference {ref:.12f}", flush=True)
        assert abs(lev @ u - ref) <= 1e-8 * max(1.0, abs(ref))
