"""Write 50-digit golden vectors for the pointwise Brinkman functions.

The formulas are transcribed from the reference's closed forms
(cuda/periodicBrinkman.h: experf :19-25, singH :31-46, regH :53-78,
diffH :85-112, vR :152-176, vF :183-208) and evaluated with mpmath at high
working precision, so cancellation in the closed forms does not reach the
50 digits stored.  The r -> 0 values come from the reference's own limit
branches (regH :73-77, diffH :105-111), also at high precision.

Two mathematical identities are recorded as well, independently of the
closed forms:
  * blob:  the 3D Fourier transform of the real-space blob
    phi_d(r) = e^{-r^2/d^2}/(pi^{3/2} d^3) [1 + (1 - e^{-a^2 d^2/4})(6 d^2 - 4 r^2)/(a^2 d^4)]
    equals [1 + (1 - e^{-a^2 d^2/4}) k^2/a^2] e^{-d^2 k^2/4}  (the symbol in vF);
  * inversion: regH(r) equals the radial inverse transform of the Fourier
    kernel phi_d(k) (I - k k^T/k^2)/(k^2 + a^2), computed by quadrature.

Run:  python tests/golden/make_pointwise_golden.py   (takes ~1 minute)
Output: tests/golden/pointwise_mp50.json
"""
import json
import os

import mpmath as mp

mp.mp.dps = 90
PI = mp.pi
SPI = mp.sqrt(mp.pi)


def experf(x, z, c=2):
    x, z = mp.mpf(x), mp.mpf(z)
    return mp.exp(c * x * z) * mp.erfc(z + x)


def singH(r, a):
    r, a = mp.mpf(r), mp.mpf(a)
    r2 = r * r
    H1 = -1 / (4 * PI * a * a * r * r2) + mp.exp(-a * r) / (4 * PI * r) * (1 + 1 / (a * r) + 1 / (a * a * r2))
    H2 = 3 / (4 * PI * a * a * r * r2 * r2) - mp.exp(-a * r) / (4 * PI * r2 * r) * (1 + 3 / (a * r) + 3 / (a * a * r2))
    return H1, H2


def regH(r, a, d):
    r, a, d = mp.mpf(r), mp.mpf(a), mp.mpf(d)
    if r == 0:
        H1 = -a * mp.erfc(a * d / 2) / (6 * PI) + (2 + mp.exp(-a * a * d * d / 4) * (a * a * d * d - 2)) / (3 * a * a * d ** 3 * PI * SPI)
        h2 = (-24 + 2 * a * a * d * d * (2 - a * a * d * d)) * mp.exp(-a * a * d * d / 4) + 24
        H2 = h2 / (60 * a * a * d ** 5 * PI * SPI) + a ** 3 / (60 * PI) * mp.erfc(a * d / 2)
        return H1, H2
    r2 = r * r
    expx2 = mp.exp(-r2 / d / d)
    Erf = mp.erf(r / d)
    om = 1 - mp.exp(-a * a * d * d / 4)
    H1 = -1 / (4 * PI * a * a * r2 * r) * Erf
    H1 += 1 / (2 * a * a * d ** 3 * PI * SPI * r2) * expx2 * (d * d + 2 * r2) * om
    H1 += 1 / (8 * a * a * PI * r2 * r) * (1 + a * r + a * a * r2) * experf(a * d / 2, -r / d)
    H1 -= 1 / (8 * a * a * PI * r2 * r) * (1 - a * r + a * a * r2) * experf(a * d / 2, r / d)
    H2 = 3 / (4 * PI * a * a * r2 * r2 * r) * Erf
    H2 -= om / (2 * a * a * d ** 3 * PI * SPI * r2 * r2) * expx2 * (3 * d * d + 2 * r2)
    H2 -= 1 / (8 * a * a * PI * r2 * r2 * r) * (3 + 3 * a * r + a * a * r2) * experf(a * d / 2, -r / d)
    H2 += 1 / (8 * a * a * PI * r2 * r2 * r) * (3 - 3 * a * r + a * a * r2) * experf(a * d / 2, r / d)
    return H1, H2


def diffH(r, a, d):
    r, a, d = mp.mpf(r), mp.mpf(a), mp.mpf(d)
    if r == 0:
        H1 = -(4 + 2 * (-2 + a * a * d * d) * mp.exp(-a * a * d * d / 4)) / (a * a * d ** 3 * 6 * PI * SPI) + a * mp.erfc(a * d / 2) / (6 * PI)
        H2 = (-(2 - a * a * d * d) * mp.exp(-a * a * d * d / 4) / (30 * d ** 3 * PI * SPI)
              - 2 * (1 - mp.exp(-a * a * d * d / 4)) / (5 * a * a * d ** 5 * PI * SPI) - a ** 3 * mp.erfc(a * d / 2) / (60 * PI))
        return H1, H2
    r2 = r * r
    expx2 = mp.exp(-r2 / d / d) * (1 - mp.exp(-a * a * d * d / 4))
    Erfc = mp.erfc(r / d)
    H1 = -1 / (4 * PI * a * a) / (r2 * r) * Erfc
    H1 -= (d * d + 2 * r2) / (2 * a * a * d ** 3 * PI * SPI * r2) * expx2
    H1 += 1 / (8 * PI * a * a * r2 * r) * (1 + a * r + a * a * r2) * experf(-a * d / 2, r / d)
    H1 += 1 / (8 * PI * a * a * r2 * r) * (1 - a * r + a * a * r2) * experf(a * d / 2, r / d)
    H2 = 3 / (4 * PI * a * a) / (r2 * r2 * r) * Erfc
    H2 += (3 * d * d + 2 * r2) / (2 * a * a * d ** 3 * PI * SPI * r2 * r2) * expx2
    H2 -= 1 / (8 * PI * a * a * r2 * r2 * r) * (3 + 3 * a * r + a * a * r2) * experf(-a * d / 2, r / d)
    H2 -= 1 / (8 * PI * a * a * r2 * r2 * r) * (3 - 3 * a * r + a * a * r2) * experf(a * d / 2, r / d)
    return H1, H2


def vR(xh, a, d, e):
    r = mp.sqrt(sum(mp.mpf(v) ** 2 for v in xh))
    H1, H2 = diffH(r, a, d)
    if e > 0:
        t1, t2 = diffH(r, a, e)
        H1, H2 = H1 - t1, H2 - t2
    A = [[mp.mpf(xh[i]) * mp.mpf(xh[j]) * H2 + (H1 if i == j else 0) for j in range(3)] for i in range(3)]
    return A


def blob_hat(k2, a, d):
    k2, a, d = mp.mpf(k2), mp.mpf(a), mp.mpf(d)
    return (1 + (1 - mp.exp(-a * a * d * d / 4)) * k2 / (a * a)) * mp.exp(-d * d * k2 / 4)


def vF(xh, a, d, k):
    k2 = sum(mp.mpf(v) ** 2 for v in k)
    kx = sum(mp.mpf(xh[i]) * mp.mpf(k[i]) for i in range(3))
    b = blob_hat(k2, a, d) * mp.cos(kx) / (k2 * (k2 + mp.mpf(a) ** 2))
    return [[b * ((k2 if i == j else 0) - mp.mpf(k[i]) * mp.mpf(k[j])) for j in range(3)] for i in range(3)]


def blob_real(r, a, d):
    r, a, d = mp.mpf(r), mp.mpf(a), mp.mpf(d)
    om = 1 - mp.exp(-a * a * d * d / 4)
    return mp.exp(-r * r / (d * d)) / (PI ** 1.5 * d ** 3) * (1 + om * (6 * d * d - 4 * r * r) / (a * a * d ** 4))


def blob_ft_quad(k, a, d):
    """3D FT of the radial blob: (4 pi/k) int_0^inf phi(r) sin(k r) r dr."""
    k = mp.mpf(k)
    f = lambda r: blob_real(r, a, d) * mp.sin(k * r) * r
    return 4 * PI / k * mp.quad(f, [0, d, 3 * d, 12 * d])


def regH_by_inversion(r, a, d):
    """H1, H2 of phi_d(k)(I - k k^T/k^2)/(k^2+a^2) by radial quadrature.

    S = G I + grad grad^T B with G^ = phi/(k^2+a^2), B^ = phi/(k^2(k^2+a^2)).
    For radial f, f(r) = (1/(2 pi^2 r)) int f^(k) k sin(kr) dk and
    B'(r)/r = -(1/(2 pi^2 r^2)) int B^ k^2 j1(kr) dk/(1)... evaluated as
    H1 = G + B'/r, H2 = (B'' - B'/r)/r^2 using spherical Bessel forms:
      G      = (1/2pi^2) int G^ k^2 j0(kr) dk
      B'/r   = -(1/2pi^2) int B^ k^3 j1(kr)/r dk
      B''-B'/r = (1/2pi^2) int B^ k^4 j2(kr) dk
    """
    r, a, d = mp.mpf(r), mp.mpf(a), mp.mpf(d)

    def j0(t):
        return mp.sin(t) / t

    def j1(t):
        return mp.sin(t) / t ** 2 - mp.cos(t) / t

    def j2(t):
        return (3 / t ** 3 - 1 / t) * mp.sin(t) - 3 * mp.cos(t) / t ** 2

    kmax = 14 / d
    pts = [mp.mpf(0)] + [kmax * i / 40 for i in range(1, 41)]
    G = mp.quad(lambda k: blob_hat(k * k, a, d) / (k * k + a * a) * k * k * j0(k * r), pts) / (2 * PI ** 2)
    Bp = -mp.quad(lambda k: blob_hat(k * k, a, d) / (k * k * (k * k + a * a)) * k ** 3 * j1(k * r) / r, pts) / (2 * PI ** 2)
    Bpp = mp.quad(lambda k: blob_hat(k * k, a, d) / (k * k * (k * k + a * a)) * k ** 4 * j2(k * r), pts) / (2 * PI ** 2)
    return G + Bp, Bpp / (r * r)


def s(v):
    return mp.nstr(v, 50)


def main():
    out = {"dps": 50, "source": "tests/golden/make_pointwise_golden.py"}
    # experf
    rows = []
    for x in [-25.0, -2.5, -0.12, -0.005, 0.005, 0.12, 2.5, 25.0]:
        for z in [0.0, 1e-3, 0.3, 1.0, 3.0, 6.0, 12.0]:
            rows.append([x, z, s(experf(x, z))])
    out["experf"] = rows
    # radial functions over (a, d) and z = r/d
    params = [(1.0, 0.05), (10.0, 0.024), (10.0, 0.1), (10.0, 1e-3), (25.0, 0.012), (25.0, 5e-4),
              (50.0, 0.01), (50.0, 0.1), (1.0, 1e-3), (3.0, 0.2), (50.0, 0.5)]
    zs = [0.0, 1e-6, 1e-4, 1e-3, 1e-2, 0.05, 0.1, 0.19, 0.21, 0.3, 0.5, 0.8, 1.0, 1.5, 1.9, 2.1, 3.0, 4.5, 6.0, 8.0]
    rad = []
    for a, d in params:
        for z in zs:
            r = z * d
            g1, g2 = regH(r, a, d)
            f1, f2 = diffH(r, a, d)
            if r > 0:
                s1, s2 = singH(r, a)
                sing = [s(s1), s(s2)]
            else:
                sing = [None, None]
            rad.append([a, d, z, s(g1), s(g2), s(f1), s(f2)] + sing)
    out["radial"] = rad  # [a, d, z, regH1, regH2, diffH1, diffH2, singH1, singH2]
    # vR tensors (column-major 9 values) incl. e = 0 and e > 0, self term
    vrs = []
    cases = [((0.01, -0.02, 0.015), 10.0, 0.024, 1e-3), ((0.0, 0.0, 0.0), 10.0, 0.024, 1e-3),
             ((0.0, 0.0, 0.0), 10.0, 0.12, 0.0), ((0.003, 0.001, -0.002), 10.0, 0.12, 0.01),
             ((0.05, 0.03, -0.04), 25.0, 0.012, 5e-4), ((2e-4, -1e-4, 3e-4), 25.0, 0.012, 5e-4),
             ((0.2, 0.1, 0.05), 1.0, 0.15, 0.0), ((4e-4, 2e-4, 1e-4), 1.0, 0.05, 1e-3),
             ((0.011, 0.0, 0.0), 50.0, 0.01, 1e-3), ((0.3, -0.2, 0.25), 2.0, 0.2, 0.05)]
    for xh, a, d, e in cases:
        A = vR(xh, a, d, e)
        vrs.append([list(xh), a, d, e, [s(A[i][j]) for j in range(3) for i in range(3)]])
    out["vR"] = vrs
    vfs = []
    fcases = [((0.1, 0.2, 0.3), 10.0, 0.024, (2 * 3.141592653589793, 0.0, 0.0)),
              ((0.4, -0.1, 0.7), 1.0, 0.15, (6.283185307179586, -12.566370614359172, 18.84955592153876)),
              ((0.25, 0.5, 0.125), 25.0, 0.012, (125.66370614359172, 62.83185307179586, -31.41592653589793)),
              ((0.9, 0.3, 0.2), 2.0, 0.2, (4.833219, 6.283185307179586, 7.853981633974483))]
    for xh, a, d, k in fcases:
        A = vF(xh, a, d, k)
        vfs.append([list(xh), a, d, list(k), [s(A[i][j]) for j in range(3) for i in range(3)]])
    out["vF"] = vfs
    # identities
    mp.mp.dps = 30
    blob = []
    for a, d in [(10.0, 0.024), (1.0, 0.1), (50.0, 0.01)]:
        for kd in [0.5, 1.5, 3.0]:
            k = kd / d
            blob.append([a, d, k, s(blob_ft_quad(k, a, d)), s(blob_hat(k * k, a, d))])
    out["blob_ft"] = blob  # [a, d, k, quadrature FT, closed-form symbol]
    inv = []
    for a, d in [(10.0, 0.1), (2.0, 0.2)]:
        for z in [0.3, 1.0, 2.0]:
            r = z * d
            q1, q2 = regH_by_inversion(r, a, d)
            mp.mp.dps = 90
            c1, c2 = regH(r, a, d)
            mp.mp.dps = 30
            inv.append([a, d, z, s(q1), s(q2), s(c1), s(c2)])
    out["inversion"] = inv  # [a, d, z, H1 quad, H2 quad, H1 closed, H2 closed]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pointwise_mp50.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
