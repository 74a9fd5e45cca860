"""The epsilon-u-Hydro and Tmunu-t outputs of MyEigen::flowVelocity4D
(MyEigen.cpp:326-800 and 801-1046, gfactor 1) through the library's host
entry points, against plain-Python restatements of the reference's loops on
synthetic fields: the hydro grid, the bilinear weights divided by the hydro
spacing, the site-0 and clamped-neighbour rules, |eps|, the rebuilt u^tau,
the vacuum and floor lines, the column signs and tau factors, the headers,
the printed Etot.  No GPU."""
import math

import numpy as np
import pytest

import ipglasma_amd as I

HBARC = 0.1973269718
# land planes of the columns after u^eta: pi tau tau, tau x, tau y, tau eta,
# xx, xy, x eta, yy, y eta, eta eta (land holds pi in T's order from plane 5)
PI_COLS = [5, 9, 10, 11, 6, 12, 13, 7, 14, 8]


def _reference_text(land, N, L, it, dtau, hx, heta, hL, deta):
    S = N * N
    a = L / N
    tau0 = it * dtau * a
    ha = hL / float(hx)
    out = ["# dummy %d etamax= %d xmax= %d ymax= %d deta= %g dx= %g dy= %g tau= %g"
           % (1, heta, hx, hx, deta, ha, ha, tau0)]
    etot = 0.0
    for ieta in range(heta):
        eta = -(heta - 1) / 2.0 * deta + deta * ieta
        for ix in range(hx):
            for iy in range(hx):
                x = -hL / 2.0 + ha * ix
                y = -hL / 2.0 + ha * iy
                line = None
                if abs(x) < L / 2.0 and abs(y) < L / 2.0:
                    xpos = int(math.floor((x + L / 2.0) / a + 0.0000000001))
                    ypos = int(math.floor((y + L / 2.0) / a + 0.0000000001))
                    xup = xpos + 1 if xpos < N - 1 else xpos
                    yup = ypos + 1 if ypos < N - 1 else ypos
                    fx = (x - (-L / 2.0 + a * xpos)) / ha
                    fy = (y - (-L / 2.0 + a * ypos)) / ha
                    p1, p2, p3, p4 = xpos * N + ypos, xup * N + ypos, xpos * N + yup, xup * N + yup
                    v = []
                    for q in range(15):
                        F = land[q]
                        g = abs if q == 0 else (lambda t: t)
                        x1 = (1 - fx) * g(F[p1]) + fx * g(F[p2]) if 0 < p1 < S and 0 < p2 < S else 0.0
                        x2 = (1 - fx) * g(F[p3]) + fx * g(F[p4]) if 0 < p3 < S and 0 < p4 < S else 0.0
                        v.append((1.0 - fy) * x1 + fy * x2)
                    v[1] = math.sqrt(1.0 + v[2] * v[2] + v[3] * v[3] + tau0 * tau0 * v[4] * v[4])
                    e = abs(HBARC * v[0] * 1.0)
                    etot += e * ha * ha * it * dtau * a
                    if e > 0.0000000001:
                        cols = [eta, x, y, e, v[1], v[2], v[3], v[4]] + [v[k] for k in PI_COLS]
                        line = " ".join("%g" % c for c in cols)
                if line is None:
                    line = " ".join("%g" % c for c in [eta, x, y, 0.0, 1.0] + [0.0] * 13)
                out.append(line)
        out.append("")
    return "\n".join(out) + "\n", etot


def _landau(N, L, rng):
    a = L / N
    x = -L / 2 + a * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    r2 = (X / (0.3 * L)) ** 2 + (Y / (0.2 * L)) ** 2
    land = np.zeros((15, N * N))
    land[0] = (np.exp(-r2) * 50 * np.where(r2 > 3.0, 1e-14, 1.0)).reshape(-1)  # vacuum far out
    land[0][5] *= -1  # |eps| is what the file takes
    land[2:5] = 0.3 * rng.normal(size=(3, N * N))
    land[1] = np.sqrt(1 + land[2] ** 2 + land[3] ** 2)
    land[5:] = rng.normal(size=(10, N * N))
    return land


@pytest.mark.parametrize("N,L,hx,heta,hL,deta", [
    (8, 4.0, 12, 3, 3.6, 0.5),    # hydro grid inside the lattice, finer than it
    (8, 4.0, 5, 1, 6.0, 1.0),     # larger than the lattice: vacuum lines outside
    (12, 6.0, 12, 2, 6.0, 0.25),  # same extent and spacing: the clamped edge
])
def test_hydro_file_matches_reference_loops(tmp_path, N, L, hx, heta, hL, deta):
    rng = np.random.default_rng(N + hx)
    land = _landau(N, L, rng)
    it, dtau = 7, 0.1
    f = tmp_path / "h.dat"
    etot = I.write_hydro_file(f, land, L, it, dtau, hx, heta, hL, deta)
    want, wetot = _reference_text(land, N, L, it, dtau, hx, heta, hL, deta)
    assert f.read_text() == want
    assert etot == pytest.approx(wetot, rel=1e-13)
    assert len(want.splitlines()[1].split()) == 18


def test_hydro_file_name_and_errors(tmp_path):
    assert I.hydro_file_name(0.4, 3) == "epsilon-u-Hydro-t0.4-3.dat"
    assert I.hydro_file_name(0.4, 3, final=True) == "epsilon-u-Hydro-TauHydro-3.dat"
    land = np.zeros((15, 16))
    with pytest.raises(I.IPGError):
        I.write_hydro_file(tmp_path / "h", land, 4.0, 1, 0.1, 0, 1, 4.0, 1.0)
    with pytest.raises(I.IPGError, match="cannot open"):
        I.write_hydro_file(tmp_path / "no" / "h", land, 4.0, 1, 0.1, 4, 1, 4.0, 1.0)


def _reference_tmunu_text(T, N, L, it, dtau, hx, heta, hL, deta):
    S = N * N
    a = L / N
    tau0 = it * dtau * a
    ha = hL / float(hx)
    small = 1e-16
    out = ["# dummy %d etamax= %d xmax= %d ymax= %d deta= %g dx= %g dy= %g" % (1, heta, hx, hx, deta, ha, ha)]
    for iy in range(hx):
        for ix in range(hx):
            x = -hL / 2.0 + ha * ix
            y = -hL / 2.0 + ha * iy
            line = None
            if abs(x) < L / 2.0 and abs(y) < L / 2.0:
                xpos = int(math.floor((x + L / 2.0) / a + 0.0000000001))
                ypos = int(math.floor((y + L / 2.0) / a + 0.0000000001))
                xup = xpos + 1 if xpos < N - 1 else xpos
                yup = ypos + 1 if ypos < N - 1 else ypos
                fx = (x - (-L / 2.0 + a * xpos)) / ha
                fy = (y - (-L / 2.0 + a * ypos)) / ha
                p1, p2, p3, p4 = xpos * N + ypos, xup * N + ypos, xpos * N + yup, xup * N + yup
                r = []
                for q in range(10):
                    F = T[q]
                    x1 = (1 - fx) * F[p1] + fx * F[p2] if 0 < p1 < S and 0 < p2 < S else 0.0
                    x2 = (1 - fx) * F[p3] + fx * F[p4] if 0 < p3 < S and 0 < p4 < S else 0.0
                    r.append((1.0 - fy) * x1 + fy * x2)
                if r[0] * 1.0 * HBARC > small:
                    cols = [r[0] * 1.0 * HBARC, r[1] * 1.0 * HBARC, r[2] * 1.0 * HBARC,
                            tau0 * tau0 * r[3] * 1.0 * HBARC, -r[4] * 1.0 * HBARC, -r[5] * 1.0 * HBARC,
                            -tau0 * r[6] * 1.0 * HBARC, -r[7] * 1.0 * HBARC, -tau0 * r[9] * 1.0 * HBARC,
                            -tau0 * r[8] * 1.0 * HBARC]
                    line = f"{ix} {iy} " + " ".join("%g" % c for c in cols)
            if line is None:
                line = f"{ix} {iy} " + " ".join("%g" % c for c in [small, small / 2.0, small / 2.0] + [0.0] * 7)
            out.append(line)
        out.append("")
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("N,L,hx,hL", [(8, 4.0, 12, 3.6), (8, 4.0, 5, 6.0), (12, 6.0, 12, 6.0)])
def test_tmunu_file_matches_reference_loops(tmp_path, N, L, hx, hL):
    rng = np.random.default_rng(N * hx)
    a = L / N
    x = -L / 2 + a * np.arange(N)
    X, Y = np.meshgrid(x, x, indexing="ij")
    T = rng.normal(size=(10, N * N))
    T[0] = (20 * np.exp(-(X ** 2 + Y ** 2) / (0.1 * L * L))).reshape(-1)
    T[0][T[0] < 1e-3] = 1e-17  # below the file's floor
    f = tmp_path / "t.dat"
    I.write_tmunu_file(f, T, L, 5, 0.1, hx, 1, hL, 1.0)
    assert f.read_text() == _reference_tmunu_text(T, N, L, 5, 0.1, hx, 1, hL, 1.0)


def _reference_jazma_text(A, B, N, L, it, dtau, hx, heta, hL, deta, Etot):
    S = N * N
    a = L / N
    ha = hL / float(hx)

    def cell(x, y):
        if not (abs(x) < L / 2.0 and abs(y) < L / 2.0):
            return None
        xpos = int(math.floor((x + L / 2.0) / a + 0.0000000001))
        ypos = int(math.floor((y + L / 2.0) / a + 0.0000000001))
        xup = xpos + 1 if xpos < N - 1 else xpos
        yup = ypos + 1 if ypos < N - 1 else ypos
        fx = (x - (-L / 2.0 + a * xpos)) / ha
        fy = (y - (-L / 2.0 + a * ypos)) / ha
        p1, p2, p3, p4 = xpos * N + ypos, xup * N + ypos, xpos * N + yup, xup * N + yup

        def f(F):
            x1 = (1 - fx) * abs(F[p1]) + fx * abs(F[p2]) if 0 < p1 < S and 0 < p2 < S else 0.0
            x2 = (1 - fx) * abs(F[p3]) + fx * abs(F[p4]) if 0 < p3 < S and 0 < p4 < S else 0.0
            return (1.0 - fy) * x1 + fy * x2
        return f(A), f(B)

    J = 0.0
    for ieta in range(heta):
        for ix in range(hx):
            for iy in range(hx):
                c = cell(-hL / 2.0 + ha * ix, -hL / 2.0 + ha * iy)
                if c is not None:
                    J += c[0] * c[1] * ha * ha * it * dtau * a
    out = ["# dummy %d etamax= %d xmax= %d ymax= %d deta= %g dx= %g dy= %g" % (1, heta, hx, hx, deta, ha, ha)]
    for ieta in range(heta):
        eta = -(heta - 1) / 2.0 * deta + deta * ieta
        for ix in range(hx):
            for iy in range(hx):
                x, y = -hL / 2.0 + ha * ix, -hL / 2.0 + ha * iy
                c = cell(x, y)
                e = 0.0 if c is None else c[0] * c[1] / J * Etot
                out.append(" ".join("%g" % v for v in [eta, x, y, e, 1.0] + [0.0] * 13))
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("N,L,hx,heta,hL", [(8, 4.0, 12, 2, 3.6), (8, 4.0, 5, 1, 6.0)])
def test_jazma_file_matches_reference_loops(tmp_path, N, L, hx, heta, hL):
    rng = np.random.default_rng(3 * N + hx)
    A = np.abs(rng.normal(size=N * N))
    B = -np.abs(rng.normal(size=N * N))  # |g2mu2| is what the file takes
    f = tmp_path / "j.dat"
    I.write_jazma_file(f, A, B, L, 6, 0.1, hx, heta, hL, 0.5, 123.4)
    assert f.read_text() == _reference_jazma_text(A, B, N, L, 6, 0.1, hx, heta, hL, 0.5, 123.4)
