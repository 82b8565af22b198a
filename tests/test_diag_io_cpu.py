"""CPU checks of the diagnostics and the CSV I/O.

Oracle diagnostics (Etot_func, postprocess.cu:54-65; Espec_func,
EnergySpec.cu:81-96) are pinned by closed forms of the Taylor-Green field and
by an identity linking the two independently written reductions.  The CSV
functions of the product library (qs_field_visual, qs_file_init, ...) are
host code: they run here, against the text the C++ stream defaults (6
significant digits, printf %g) give, and against the reference's parsing
rules (friction.cu:62-97)."""
import os

import numpy as np
import pytest

import oracle_py as O
import qsolver as Q

TWO_PI = 2 * np.pi


def taylor_green(N):
    m = O.Mesh(N, N, TWO_PI)
    X, Y = m.grid()
    w = -2.0 * np.cos(X) * np.cos(Y)
    return m, w, m.fwd(w)


def test_etot_closed_form():
    """u = cos x sin y, v = -sin x cos y: sum (u^2+v^2)/2 dx dy = pi^2, i.e.
    Etot / (Lx Ly) = 1/4 (postprocess.cu:160-161)."""
    m, _, _ = taylor_green(16)
    X, Y = m.grid()
    u, v = np.cos(X) * np.sin(Y), -np.sin(X) * np.cos(Y)
    E = m.Etot(u, v)
    assert E == pytest.approx(np.pi ** 2, rel=1e-14)
    assert E / (m.Lx * m.Ly) == pytest.approx(0.25, rel=1e-14)


def test_espec_closed_form():
    """omega = -2 cos x cos y has |K| = sqrt(2): with dk = 1 all of it lies in
    shell [1, 2).  Its two half-spectrum bins (1, +-1) carry |u|^2 + |v|^2 =
    N^4/8 each, so E_1 = (N^4/4) / (2 dk N^4) = 1/8 and every other shell is 0."""
    N = 16
    m, _, wt = taylor_green(N)
    ut, vt = m.vel(wt)
    E = m.Espec_shells(ut, vt, 1.0, 8.0)
    assert len(E) == 8
    assert E[1] == pytest.approx(1.0 / 8.0, rel=1e-13)
    assert np.all(np.abs(np.delete(E, 1)) < 1e-28)
    assert m.Espec(1.0, ut, vt, 1.0) == E[1]


def test_espec_shell_edges_follow_the_accumulated_k():
    """The driver's k is accumulated by repeated addition (EnergySpec.cu:
    198-203).  With its box L = 66 pi, k_n so accumulated exceeds the axis
    wavenumber n dk from n = 273 on, so the mode kx = 274 dk lies in shell 273,
    not in shell 274 as a recomputed k_n = n dk would place it."""
    N, L = 1024, 66 * np.pi
    m = O.Mesh(N, N, L)
    dk = 2 * np.pi / L
    wt = np.zeros(m.spec_shape, complex)
    wt[0, 274] = N * N
    ut, vt = m.vel(wt)
    E = m.Espec_shells(ut, vt, dk, 280 * dk)
    assert np.flatnonzero(E).tolist() == [273]


def test_espec_sums_to_half_the_kinetic_energy():
    """For a vorticity field without kx = 0 and kx = Nx/2 content every
    conjugate pair appears once in the half spectrum, so over all shells
    sum_n E_n dk = Etot / (2 Lx Ly): the spectrum (EnergySpec.cu) and the
    physical-space energy (postprocess.cu) are independent reductions."""
    N = 32
    m = O.Mesh(N, N, TWO_PI)
    rng = np.random.default_rng(4)
    w = rng.standard_normal(m.phys_shape)
    wt = m.fwd(w)
    wt[:, 0] = 0.0
    wt[:, -1] = 0.0
    ut, vt = m.vel(wt)
    u, v = m.bwd(ut.copy()), m.bwd(vt.copy())
    E = m.Espec_shells(ut, vt, 1.0, 2.0 * N)
    assert np.sum(E) * 1.0 == pytest.approx(m.Etot(u, v) / (2 * m.Lx * m.Ly), rel=1e-12)


# ---- CSV I/O of the product library (host code) ------------------------------

def csv_text(a):
    return "".join("".join("%g," % x for x in row) + "\n" for row in a)


def edge_values(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) * 10.0 ** rng.integers(-8, 9, shape)
    flat = a.reshape(-1)
    flat[:6] = [0.0, -0.0, 1e-5, 123456789.0, -1.5e-300, 0.1]
    return a


def test_field_visual_writes_the_reference_text(tmp_path):
    a = edge_values((5, 7), 1)
    p = tmp_path / "f.csv"
    Q.field_visual(a, str(p))
    assert p.read_text() == csv_text(a)


def test_wave_and_spec_visual_text(tmp_path):
    a = edge_values((4, 3), 2)
    z = a + 1j * edge_values((4, 3), 3)
    Q.wave_visual(a, str(tmp_path / "k.txt"))
    Q.spec_visual(z, str(tmp_path / "z.txt"))
    assert (tmp_path / "k.txt").read_text() == csv_text(a)
    want = "".join("".join("(%g,%g)," % (c.real, c.imag) for c in row) + "\n" for row in z)
    assert (tmp_path / "z.txt").read_text() == want


def test_coord_writes_single_precision_coordinates(tmp_path):
    x, y = tmp_path / "x.csv", tmp_path / "y.csv"
    dx = 66 * np.pi / 16
    Q.coord(dx, dx, 16, 4, str(x), str(y))
    X = np.float32(dx * np.arange(16)).astype(np.float64)
    assert x.read_text() == csv_text(np.tile(X, (4, 1)))
    assert y.read_text() == csv_text(np.tile(np.float32(dx * np.arange(4))[:, None].astype(np.float64), (1, 16)))


class _HostMesh:
    """A Mesh handle for the host-only qs_file_init (reads Nx, Ny only)."""

    def __init__(self, Nx, Ny):
        self.s = Q.MeshStruct()
        self.s.Nx, self.s.Ny, self.s.Nxh = Nx, Ny, Nx // 2 + 1


def file_init(path, Nx, Ny):
    import ctypes
    hm = _HostMesh(Nx, Ny)
    out = np.full((Ny, Nx), np.pi)
    Q.call("qs_file_init", os.fsencode(str(path)), out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(hm.s))
    return out


def test_file_init_reads_what_field_visual_wrote(tmp_path):
    a = edge_values((6, 8), 5)
    p = tmp_path / "w.csv"
    Q.field_visual(a, str(p))
    got = file_init(p, 8, 6)
    want = np.array([[float("%g" % x) for x in row] for row in a])
    assert np.array_equal(got, want)


def test_file_init_parsing_rules(tmp_path):
    """getline(ss, str, ','): a trailing comma adds no value, an inner empty
    token is one; string2num of text that does not convert is 0."""
    p = tmp_path / "a.csv"
    p.write_text("1,2.5,3e2\n-4, 5,junk\n")
    assert np.array_equal(file_init(p, 3, 2), [[1, 2.5, 300], [-4, 5, 0]])
    p.write_text("1,,3,\n4,5,6,\n")
    assert np.array_equal(file_init(p, 3, 2), [[1, 0, 3], [4, 5, 6]])


@pytest.mark.parametrize("text,Nx,Ny", [("1,2,\n3,4,\n", 3, 2), ("1,2,\n", 2, 2), ("1,2,\n3,\n", 2, 2)])
def test_file_init_rejects_mismatched_sizes(tmp_path, text, Nx, Ny):
    p = tmp_path / "b.csv"
    p.write_text(text)
    with pytest.raises(Q.QsolverError) as e:
        file_init(p, Nx, Ny)
    assert e.value.status == Q.QS_ERR_INVALID_ARG


def test_file_init_missing_file(tmp_path):
    with pytest.raises(Q.QsolverError) as e:
        file_init(tmp_path / "none.csv", 2, 2)
    assert e.value.status == Q.QS_ERR_IO
