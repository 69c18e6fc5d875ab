"""Generate the high-precision golden fixtures of tests/golden/ (run once, committed).

Independent of both the oracle and the GPU library: mpmath (40 digits) and sympy
(exact) evaluate the closed forms.  Data files copied verbatim from the reference's
own test data live next to this script: spherical_hankel.txt
(crates/treams-core/references/), degrees.txt and harmonics.txt (formal/golden/).
The end-to-end numbers of the reference's documented examples are in
reference_examples.json (inputs restated from docs/examples/*.py, outputs from
docs/examples/output/*.txt).

    python tests/golden/make_fixtures.py
"""
import json
import os

import mpmath as mp
from sympy.physics.wigner import wigner_3j

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 40


def c2(z):
    z = mp.mpc(z)
    return [float(z.real), float(z.imag)]


def sph_h(p, x):
    x = mp.mpf(x)
    return mp.sqrt(mp.pi / (2 * x)) * (mp.besselj(p + mp.mpf(1) / 2, x) + 1j * mp.bessely(p + mp.mpf(1) / 2, x))


def sph_j(p, z):
    z = mp.mpc(z)
    return mp.sqrt(mp.pi / (2 * z)) * mp.besselj(p + mp.mpf(1) / 2, z)


def hankel_table():
    xs = [0.4, 0.455, 0.5, 0.65, 1.0, 1.3, 2.6, 5.0, 13.0, 40.0]
    return {"x": xs, "pmax": 13, "values": [[c2(sph_h(p, x)) for p in range(14)] for x in xs]}


def jn_table():
    zs = [0.3, complex(0.45, 0.1), 0.7, 1.3, complex(3.0, 0.5), complex(5.66, 22.6), 12.0, complex(0.2, -0.3)]
    # zeros of j_0 (k pi) and of j_1 (4.4934..., 7.7253...), where a sequence scaled by
    # one of them alone loses its accuracy, and a complex point next to pi
    zs += [float(mp.pi), float(2 * mp.pi), float(3 * mp.pi), 4.493409457909064, 7.725251836937707,
           complex(float(mp.pi), 1e-3)]
    return {"z": [[z.real, z.imag] if isinstance(z, complex) else [z, 0.0] for z in zs], "pmax": 13,
            "values": [[c2(sph_j(p, z)) for p in range(14)] for z in zs]}


def w3j_table():
    rows = []
    for j1 in range(0, 5):
        for j2 in range(0, 5):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        m3 = -m1 - m2
                        if abs(m3) > j3:
                            continue
                        rows.append([j1, j2, j3, m1, m2, m3, float(wigner_3j(j1, j2, j3, m1, m2, m3))])
    for (j1, j2, m1, m2) in [(6, 6, 0, 0), (6, 6, 3, -5), (6, 5, -6, 2), (12, 6, 4, -1), (6, 6, 6, 6), (9, 7, -2, 5)]:
        for j3 in range(max(abs(j1 - j2), abs(m1 + m2)), j1 + j2 + 1):
            rows.append([j1, j2, j3, m1, m2, -m1 - m2, float(wigner_3j(j1, j2, j3, m1, m2, -m1 - m2))])
    return rows


def riccati_d(f, n, z):
    """[z f_n(z)]' = z f_{n-1}(z) - n f_n(z)."""
    return z * f(n - 1, z) - n * f(n, z)


def mie_bh(n, x, eps):
    """Bohren-Huffman a_n, b_n of a nonmagnetic sphere (relative index m) in vacuum."""
    x = mp.mpf(x)
    m = mp.sqrt(mp.mpc(eps))
    mx = m * x
    jx = sph_j(n, x)
    jmx = sph_j(n, mx)
    hx = sph_h(n, x)
    djx = riccati_d(sph_j, n, x)
    dhx = riccati_d(lambda p, z: sph_h(p, z), n, x)
    djmx = riccati_d(sph_j, n, mx)
    a = (m ** 2 * jmx * djx - jx * djmx) / (m ** 2 * jmx * dhx - hx * djmx)
    b = (jmx * djx - jx * djmx) / (jmx * dhx - hx * djmx)
    return a, b


def mie_table():
    cases = [(0.3, complex(4, 0.1)), (0.455, complex(6, 0.2)), (1.1, complex(2.5, 0)), (2.0, complex(4, 0.1)),
             (0.26, complex(16, 0.5)), (float(mp.pi), complex(4, 0.1)), (4.493409457909064, complex(2.5, 0))]
    out = []
    for x, eps in cases:
        for n in range(1, 7):
            a, b = mie_bh(n, x, eps)
            out.append({"x": x, "eps": [eps.real, eps.imag], "n": n, "a": c2(a), "b": c2(b)})
    return out


REFERENCE = "/root/reference"


def cluster_example_fields():
    """The 7 x 7 total-field intensity maps (y = 0, x and z on linspace(-300, 300, 7)
    nm) of the local expansion (NaN inside a sphere) and of the expansion about the
    origin with lmax 6 (NaN within radius 260 nm), and that expansion's cross sections,
    of docs/examples/cluster.py, read from docs/examples/output/cluster.txt into
    reference_examples.json."""
    path = os.path.join(REFERENCE, "docs", "examples", "output", "cluster.txt")
    with open(path) as f:
        lines = f.read().splitlines()
    def grid_map(label):
        start = lines.index(f"Intensity at y = 0, {label} expansion (rows: x, columns: z)") + 1
        return [[None if v == "nan" else float(v) for v in line.split()] for line in lines[start:start + 7]]

    rows = grid_map("local")
    out = os.path.join(HERE, "reference_examples.json")
    with open(out) as f:
        data = json.load(f)
    glob = next(line.split() for line in lines if line.split()[:1] == ["global"])
    data["cluster_example"]["expected_global"] = {
        "lmax": 6, "scattering": float(glob[1]), "extinction": float(glob[2]), "absorption": float(glob[3]),
        "significant_digits": 6}
    data["cluster_example"]["intensity_local"] = {
        "grid": "x, z = linspace(-300, 300, 7) nm, y = 0; rows x, columns z; null inside a sphere",
        "values": rows, "significant_digits": 6}
    rot = next(line.split() for line in lines if line.split()[:1] == ["rotated"])
    data["cluster_example"]["expected_rotated"] = {
        "lmax": 6, "euler": [0.0, 1.5707963267948966, 0.0], "direction": [0, 0, -1],
        "scattering": float(rot[1]), "extinction": float(rot[2]), "absorption": float(rot[3]),
        "significant_digits": 6}
    data["cluster_example"]["intensity_rotated"] = {
        "grid": "x, z = linspace(-300, 300, 7) nm, y = 0; rows x, columns z; null within radius 260 nm "
                "(the expansion about the origin rotated by z-y-z Euler angles (0, pi/2, 0), lit along -z)",
        "values": grid_map("rotated"), "significant_digits": 6}
    data["cluster_example"]["intensity_global"] = {
        "grid": "x, z = linspace(-300, 300, 7) nm, y = 0; rows x, columns z; null within radius 260 nm "
                "(the expansion about the origin, lmax 6)",
        "values": grid_map("global"), "significant_digits": 6}
    with open(out, "w") as f:
        json.dump(data, f, indent=1)


def main():
    cluster_example_fields()
    data = {
        "hankel": hankel_table(),
        "jn": jn_table(),
        "wigner3j": w3j_table(),
        "mie_bohren_huffman": mie_table(),
    }
    with open(os.path.join(HERE, "special_functions.json"), "w") as f:
        json.dump(data, f, indent=0)


if __name__ == "__main__":
    main()
