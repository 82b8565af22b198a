"""GPU diagnostics (qs_Etot_func, qs_Espec_func, qs_Espec_shells) against the
oracle, and the solver-level snapshot / restart I/O: CSV snapshot and CSV
initialisation (friction.cu:163-185, :262-279) and exact binary checkpoints."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qsolver as Q  # noqa: E402
import oracle_py as O  # noqa: E402
from qsolver import synth  # noqa: E402

P = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, D_Q=0.05, a_Q=1.0, c_Q=1.0, dt=1e-3)


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def csv_text(a):
    return "".join("".join("%g," % x for x in row) + "\n" for row in a)


@pytest.mark.parametrize("N", [64, 256])
def test_etot_vs_oracle(N):
    print(f"[gpu-test] Etot N={N}", flush=True)
    rng = np.random.default_rng(N)
    u, v = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    m = Q.Mesh(N, N, 2 * np.pi)
    got = m.Etot(cu(u), cu(v))
    want = O.Mesh(N, N, 2 * np.pi).Etot(u, v)
    assert got == pytest.approx(want, rel=1e-12)
    # deterministic: the same bits on a second call
    assert m.Etot(cu(u), cu(v)) == got


@pytest.mark.parametrize("N,L", [(64, 2 * np.pi), (256, 66 * np.pi)])
def test_espec_shells_vs_oracle(N, L):
    print(f"[gpu-test] Espec N={N}", flush=True)
    om = O.Mesh(N, N, L)
    rng = np.random.default_rng(7)
    wt = om.fwd(rng.standard_normal((N, N)))
    ut, vt = om.vel(wt)
    dk = 2 * np.pi / L
    kmax = 180 * dk
    want = om.Espec_shells(ut, vt, dk, kmax)
    m = Q.Mesh(N, N, L)
    got = m.Espec_shells(cu(ut), cu(vt), dk, kmax).cpu().numpy()
    assert got.shape == want.shape
    nz = want != 0
    assert np.array_equal(got != 0, nz)
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) < 1e-12
    k = 5 * dk
    assert m.Espec(k, cu(ut), cu(vt), dk) == pytest.approx(om.Espec(k, ut, vt, dk), rel=1e-12)


def test_espec_shell_edge_on_gpu():
    """The accumulated-k shell edges of EnergySpec.cu:198-203 at the driver's
    L = 66 pi: the axis mode kx = 274 dk falls in shell 273 (the oracle's
    membership, pinned on the CPU)."""
    print("[gpu-test] Espec shell edge", flush=True)
    N, L = 1024, 66 * np.pi
    om = O.Mesh(N, N, L)
    dk = 2 * np.pi / L
    wt = np.zeros(om.spec_shape, complex)
    wt[0, 274] = N * N
    wt[5, 100] = 0.5 * N * N
    ut, vt = om.vel(wt)
    m = Q.Mesh(N, N, L)
    got = m.Espec_shells(cu(ut), cu(vt), dk, 280 * dk).cpu().numpy()
    want = om.Espec_shells(ut, vt, dk, 280 * dk)
    assert np.flatnonzero(got).tolist() == np.flatnonzero(want).tolist()
    assert 273 in np.flatnonzero(got).tolist()
    assert rel(got, want) < 1e-12


def initial(N, seed):
    q11, q12 = synth.uniform_q(N, N, seed=seed)
    mesh = Q.Mesh(N, N, 2 * np.pi)
    st = [mesh.spec(), mesh.spec(), mesh.spec()]
    mesh.FwdTrans(st[1], cu(q11))
    mesh.FwdTrans(st[2], cu(q12))
    return mesh, st


def test_binary_checkpoint_resumes_bit_for_bit(tmp_path):
    print("[gpu-test] save/load", flush=True)
    N = 64
    mesh, st = initial(N, 3)
    Ra = cu(synth.activity_field(N, N, seed=4))
    a = Q.Solver(mesh, Q.params(**P))
    a.set_state(*st)
    a.set_activity(Ra)
    a.step(10)
    path = str(tmp_path / "run.qsnap")
    a.save(path, step=10)
    a.step(7)
    want = [t.clone() for t in a.get_state()]
    b = Q.Solver(mesh, Q.params(**P))
    step, p = b.load(path)
    assert step == 10 and p.dt == P["dt"] and p.alpha == P["alpha"]
    b.step(7)
    for x, y in zip(b.get_state(), want):
        assert torch.equal(x, y)
    # a snapshot without activity clears the loaded solver's
    a.set_activity(None)
    a.save(path, step=0)
    b.load(path)
    a.step(3)
    b.step(3)
    for x, y in zip(b.get_state(), a.get_state()):
        assert torch.equal(x, y)
    a.close()
    b.close()


def test_binary_checkpoint_errors(tmp_path):
    print("[gpu-test] save/load errors", flush=True)
    mesh, st = initial(32, 5)
    a = Q.Solver(mesh, Q.params(**P))
    a.set_state(*st)
    path = tmp_path / "s.qsnap"
    a.save(str(path), 1)
    other = Q.Solver(Q.Mesh(64, 64, 2 * np.pi), Q.params(**P))
    with pytest.raises(Q.QsolverError) as e:
        other.load(str(path))
    assert e.value.status == Q.QS_ERR_INVALID_ARG
    raw = bytearray(path.read_bytes())
    raw[200] ^= 0x01
    path.write_bytes(bytes(raw))
    with pytest.raises(Q.QsolverError) as e:
        a.load(str(path))
    assert e.value.status == Q.QS_ERR_IO
    with pytest.raises(Q.QsolverError) as e:
        a.load(str(tmp_path / "missing.qsnap"))
    assert e.value.status == Q.QS_ERR_IO
    a.close()
    other.close()


def test_snapshot_csv_and_csv_init(tmp_path):
    """snapshot_csv writes <prefix><step>{w,r1,r2,S}.csv as friction.cu:262-279
    does: BwdTrans of each state spectrum (whose column-0 symmetrisation stays
    in the state) and S = 2 sqrt(r1^2 + r2^2); init_csv reads such files back
    through file_init, FwdTrans and the BwdTrans round trip (:163-185)."""
    print("[gpu-test] snapshot csv", flush=True)
    N = 64
    mesh, st = initial(N, 6)
    om = O.Mesh(N, N, 2 * np.pi)
    rng = np.random.default_rng(8)
    w0 = om.fwd(0.3 * rng.standard_normal((N, N)))
    w0[3, 0] += 0.25 + 0.5j   # a non-Hermitian column 0, symmetrised by the snapshot
    st[0] = cu(w0)
    s = Q.Solver(mesh, Q.params(**P))
    s.set_state(*st)
    prefix = str(tmp_path / "snap_")
    s.snapshot_csv(prefix, 1000)
    # the files hold the reference text of the library's own BwdTrans
    ref = []
    for name, t in zip(("w", "r1", "r2"), st):
        spec = t.clone()
        f = mesh.phys()
        mesh.BwdTrans(f, spec)
        ref.append(f.cpu().numpy())
        assert (tmp_path / f"snap_1000{name}.csv").read_text() == csv_text(ref[-1])
    S = 2.0 * np.sqrt(ref[1] ** 2 + ref[2] ** 2)
    got_S = np.array([[float(x) for x in line.split(",")[:-1]]
                      for line in (tmp_path / "snap_1000S.csv").read_text().splitlines()])
    assert np.max(np.abs(got_S - S)) <= 5e-6 * np.max(np.abs(S))
    # ... which agrees with the oracle's BwdTrans
    assert rel(ref[0], om.bwd(w0.copy())) < 1e-12
    # the state's column 0 was symmetrised in place
    w_after = s.get_state()[0].cpu().numpy()
    col = w_after[:, 0]
    assert np.array_equal(col[1:N // 2], np.conj(col[N - 1:N // 2:-1]))
    # CSV initialisation from the snapshot files
    s2 = Q.Solver(mesh, Q.params(**P))
    s2.init_csv(prefix + "1000w.csv", prefix + "1000r1.csv", prefix + "1000r2.csv")
    got = [t.cpu().numpy() for t in s2.get_state()]
    for name, g in zip(("w", "r1", "r2"), got):
        f = mesh.file_init(prefix + f"1000{name}.csv")
        want = om.fwd(f)
        om.bwd(want)          # the round trip symmetrises column 0 of want in place
        assert rel(g, want) < 1e-12, name
    s.close()
    s2.close()
