"""Generate the golden fixtures in tests/golden/*.npz from the compiled
reference (oracle/_ref/ref_driver, built from /root/reference by
oracle/Makefile.ref). Run in the build container only:

    make -f oracle/Makefile.ref -j8 && python tests/golden/make_fixtures.py

Each fixture holds the model parameters, the reference's CSR (for small
models) and the reference's outputs for the stated queries. The models are
written with tbtk_amd.lattice.write_amplitudes and read back by the driver
through Model << HoppingAmplitude, so basis numbering and CSR assembly are the
reference's own.
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tbtk_amd import lattice as L  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
TMP = tempfile.mkdtemp(prefix="tbtk_fix_")


def idx(t):
    return ",".join(str(v) for v in t)


def ref_csr(model, name):
    path = os.path.join(TMP, name + ".bin")
    L.write_amplitudes(model, path)
    prefix = os.path.join(TMP, name)
    O.run_ref(mode="csr", model=path, out=prefix)
    rp = np.fromfile(prefix + ".rowptr", np.uint32)
    col = np.fromfile(prefix + ".col", np.uint32)
    val = np.fromfile(prefix + ".val", np.complex128)
    return path, (rp, col, val)


def ref_coef(path, scale, M, broadening, frm, tos):
    out = os.path.join(TMP, "coef.bin")
    O.run_ref(mode="coef", model=path, scale=scale, M=M, broadening=broadening,
              **{"from": idx(frm)}, to=";".join(idx(t) for t in tos), out=out)
    return np.fromfile(out, np.complex128).reshape(len(tos), M)


def save(name, meta, **arrays):
    np.savez_compressed(os.path.join(OUT, name + ".npz"),
                        meta=np.array(json.dumps(meta)), **arrays)
    print("wrote", name, {k: v.shape for k, v in arrays.items()})


def csr_digest(rp, col, val) -> str:
    """SHA-256 of a CSR's bytes: rowptr (u32), col (u32), val (complex128),
    little endian, concatenated."""
    import hashlib
    h = hashlib.sha256()
    for a, dt in ((rp, "<u4"), (col, "<u4"), (val, "<c16")):
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


def moments_fixture(name, model, scale, M, frm, tos, broadenings=(0.0,),
                    keep_csr=True):
    """Moments of the reference for the given queries; the reference's CSR
    itself for small models, else its SHA-256 digest (csr_sha256)."""
    path, csr = ref_csr(model, name)
    arrays = {}
    if keep_csr:
        arrays.update(rowptr=csr[0], col=csr[1], val=csr[2])
    else:
        arrays["csr_sha256"] = np.array(csr_digest(*csr))
    for b in broadenings:
        arrays[f"coef_b{b:g}"] = ref_coef(path, scale, M, b, frm, tos)
    meta = dict(model=model.name, params=model.params, shape=list(model.shape),
                scale=scale, M=M, frm=list(frm), to=[list(t) for t in tos],
                broadenings=list(broadenings))
    save(name, meta, **arrays)
    return path


def thermo_fixture(p1):
    """Thermodynamic-equilibrium properties (Fermi-Dirac weights) on the
    30x30 spin model (amplitude file p1): Density, Magnetization and
    SpinPolarizedLDOS in the Custom and the Ranges formats, expectation
    values; T = 300 K and T = 0, mu = 0.3."""
    th = {}
    base = dict(model=p1, scale=10.0, M=400, broadening=1e-6, window="-9.5,9.5,300",
                lookup=1, mu=0.3)
    for T in (300.0, 0.0):
        tag = f"T{T:g}"
        out = os.path.join(TMP, "th.bin")
        O.run_ref(mode="density", patterns="15,15,+;3,4,*;0,0,1", T=T, out=out, **base)
        th[f"density_custom_{tag}"] = np.fromfile(out, np.float64)
        O.run_ref(mode="density", pattern="x,15,+", ranges="4,1,2", T=T, out=out, **base)
        th[f"density_ranges_{tag}"] = np.fromfile(out, np.float64)
        O.run_ref(mode="mag", patterns="15,15,s;2,3,s", T=T, out=out, **base)
        th[f"mag_custom_{tag}"] = np.fromfile(out, np.complex128)
        O.run_ref(mode="mag", pattern="x,15,s", ranges="3,1,2", T=T, out=out, **base)
        th[f"mag_ranges_{tag}"] = np.fromfile(out, np.complex128)
        O.run_ref(mode="expval", to="15,15,0", **{"from": "16,15,0"}, T=T, out=out, **base)
        th[f"expval_offdiag_{tag}"] = np.fromfile(out, np.complex128)
        O.run_ref(mode="expval", to="15,15,1", **{"from": "15,15,1"}, T=T, out=out, **base)
        th[f"expval_diag_{tag}"] = np.fromfile(out, np.complex128)
    out = os.path.join(TMP, "th.bin")
    O.run_ref(mode="splodos", patterns="15,15,s", T=300.0, out=out, **base)
    th["splodos_custom"] = np.fromfile(out, np.complex128)
    O.run_ref(mode="splodos", pattern="x,14,s", ranges="2,1,2", T=300.0, out=out, **base)
    th["splodos_ranges"] = np.fromfile(out, np.complex128)
    save("square30_spin_thermo",
         dict(scale=10.0, M=400, broadening=1e-6, window=[-9.5, 9.5, 300], lookup=1,
              mu=0.3, T=[300.0, 0.0],
              density_custom="15,15,+;3,4,*;0,0,1", density_ranges=["x,15,+", "4,1,2"],
              mag_custom="15,15,s;2,3,s", mag_ranges=["x,15,s", "3,1,2"],
              expval=[["15,15,0", "16,15,0"], ["15,15,1", "15,15,1"]],
              splodos_custom="15,15,s", splodos_ranges=["x,14,s", "2,1,2"]),
         **th)



def json_fixture(p1, pp):
    """The reference's own JSON serializations (serialize(Serializable::
    Mode::JSON)) of the properties tests/cpp/pe_outputs.cpp computes with the
    same parameters: LDOS (Custom, Ranges), a NonPrincipal Green's function,
    Density (Custom, Ranges, T = 300 K, mu = 0.3) and sampleDOS. Written as
    tests/golden/json/<name>.json."""
    jdir = os.path.join(OUT, "json")
    os.makedirs(jdir, exist_ok=True)
    base = dict(model=p1, scale=10.0, M=400, broadening=1e-6, window="-9.5,9.5,300", lookup=1)
    out = os.path.join(TMP, "j.bin")
    cases = [
        ("ldos_custom", dict(mode="ldos", patterns="15,15,+;3,4,*", **base)),
        ("ldos_ranges", dict(mode="ldosr", pattern="x,15,+", ranges="2,1,2", **base)),
        ("gf_NonPrincipal", dict(mode="gf", type="NonPrincipal", to="15,15,0;16,15,0",
                                 **{"from": "15,15,0"}, **base)),
        ("density_custom_T300", dict(mode="density", patterns="15,15,+;3,4,*;0,0,1", T=300.0,
                                     mu=0.3, **base)),
        ("density_ranges_T300", dict(mode="density", pattern="x,15,+", ranges="4,1,2", T=300.0,
                                     mu=0.3, **base)),
        ("sampledos", dict(mode="sampledos", model=pp, scale=10.0, M=200, broadening=1e-6,
                           window="-9,9,200", lookup=1, n=7, seed=3)),
    ]
    for name, kw in cases:
        O.run_ref(out=out, json=os.path.join(jdir, name + ".json"), **kw)


def callers_fixture():
    """The reference's own callers of the path, which leave both GPU flags
    false and use the lookup table.

    doc_example_chebyshev.txt: stdout of doc/examples/Solver/
    ChebyshevExpander.cpp compiled as written against the reference library
    (oracle/_ref/doc_example_ref, built by oracle/Makefile.ref).

    basic_chebyshev_template.npz: the LDOS of Templates/BasicChebyshev/src/
    main.cpp:91-116 (100x100 with spin, mu = -1, t = 1, open edges; s = 5,
    500 coefficients, 1000 energies, {_a_, 50, IDX_SUM_ALL}). As written the
    template sets the window [-5, 5] at s = 5, which the reference rejects
    (exit 1), so the fixture uses [-4.99, 4.99]. Stored as the SHA-256 of
    the LDOS bytes (float64, little endian) and the rows x = 0, 50, 99."""
    import hashlib
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "doc_example_ref")
    r = subprocess.run([exe], check=True, capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    with open(os.path.join(OUT, "doc_example_chebyshev.txt"), "w") as f:
        f.write(r.stdout)
    print("wrote doc_example_chebyshev.txt")
    m = L.square_spin(100, 100, t=1.0, J=0.0, mu=-1.0)
    path = os.path.join(TMP, "template.bin")
    L.write_amplitudes(m, path)
    out = os.path.join(TMP, "template_ldos.bin")
    O.run_ref(mode="ldos", model=path, scale=5.0, M=500, broadening=1e-6,
              window="-4.99,4.99,1000", lookup=1, patterns="*,50,+", out=out)
    ldos = np.fromfile(out, np.float64)
    assert ldos.size == 100 * 1000, ldos.size
    rows = ldos.reshape(100, 1000)[[0, 50, 99]]
    # The pattern form of calculateGreensFunction with 900 from-indices
    # (Lib/src/PropertyExtractor/ChebyshevExpander.cpp:107-232): G((15,15,0),
    # (x,y,0)) for every x, y on the 30x30 spin lattice, Retarded, M = 400,
    # lookup table; 900 x 300 values as SHA-256 plus the first 4 from-indices.
    p1 = os.path.join(TMP, "sq30.bin")
    L.write_amplitudes(L.config_plain_small(), p1)
    out = os.path.join(TMP, "gfpat.bin")
    O.run_ref(mode="gfpat", model=p1, scale=10.0, M=400, broadening=1e-6,
              window="-9.5,9.5,300", lookup=1, type="Retarded", to="15,15,0", **{"from": "*,*,0"},
              out=out)
    gp = np.fromfile(out, np.complex128)
    assert gp.size == 900 * 300, gp.size
    save("square30_spin_gfpattern900",
         dict(scale=10.0, M=400, broadening=1e-6, window=[-9.5, 9.5, 300], lookup=1,
              type="Retarded", to="15,15,0", frm="*,*,0"),
         sha256=np.array(hashlib.sha256(gp.astype("<c16").tobytes()).hexdigest()),
         first=gp.reshape(900, 300)[:4])
    save("basic_chebyshev_template",
         dict(model="square_spin(100, 100, t=1, J=0, mu=-1)", scale=5.0, M=500,
              broadening=1e-6, window=[-4.99, 4.99, 1000], pattern="_a_,50,IDX_SUM_ALL",
              rows=[0, 50, 99]),
         sha256=np.array(hashlib.sha256(ldos.astype("<f8").tobytes()).hexdigest()),
         rows=rows)


def main(only=()):
    """Write every fixture, or only those named in `only` ("callers" alone
    writes just the reference-caller fixtures)."""
    if tuple(only) == ("callers",):
        callers_fixture()
        return
    if only:
        global save
        _save = save

        def save(name, meta, **arrays):  # noqa: F811
            if name in only:
                _save(name, meta, **arrays)
    if not O.ref_available():
        sys.exit("build oracle/_ref/ref_driver first (make -f oracle/Makefile.ref)")

    # Known-answer model of Test/include/TBTK/Test/Solver/ChebyshevExpander.h:184-277
    moments_fixture("periodic5", L.periodic_square(5), 10.0, 100, (0, 0),
                    [(0, 0), (1, 0), (2, 3)])

    # 30x30 with spin (1,800 states), site (15,15,s)
    m1 = L.config_plain_small()
    p1 = moments_fixture("square30_spin", m1, 10.0, 1000, (15, 15, 0),
                         [(15, 15, 0), (16, 15, 0), (15, 15, 1), (0, 0, 0), (15, 15, 0)],
                         broadenings=(0.0, 1e-6))
    # LDOS at (15,15,IDX_SUM_ALL) on [-9.99, 9.99] x 1000 (the literal window
    # [-10, 10] is rejected by the reference at scale factor 10).
    ld = {}
    for lk in (0, 1):
        out = os.path.join(TMP, "ldos.bin")
        O.run_ref(mode="ldos", model=p1, scale=10.0, M=1000, broadening=1e-6,
                  window="-9.99,9.99,1000", lookup=lk, patterns="15,15,+", out=out)
        ld[f"ldos_lookup{lk}"] = np.fromfile(out, np.float64)
    # Green's functions of the four types from one from-index, two to-indices
    for t, name in O.TYPE_NAMES.items():
        out = os.path.join(TMP, "gf.bin")
        O.run_ref(mode="gf", model=p1, scale=10.0, M=400, broadening=1e-6,
                  window="-9.5,9.5,300", lookup=1, type=name,
                  **{"from": "15,15,0"}, to="15,15,0;16,15,0", out=out)
        ld[f"gf_{name}"] = np.fromfile(out, np.complex128).reshape(2, 300)
    save("square30_spin_ldos",
         dict(scale=10.0, M=1000, window=[-9.99, 9.99, 1000], pattern="15,15,+",
              broadening=1e-6, gf=dict(M=400, window=[-9.5, 9.5, 300],
                                      frm=[15, 15, 0], to=[[15, 15, 0], [16, 15, 0]])),
         **ld)

    thermo_fixture(p1)

    # Green's functions from given coefficients (no model): c = {1,2,3} on
    # Range(-5,5,10) as in Test/.../Solver/ChebyshevExpander.h:371-565, and a
    # random complex c of length 50 on Range(-8.7, 9.1, 37).
    gfa = {}
    rng = np.random.default_rng(11)
    cases = {
        "c123": (np.array([1, 2, 3], dtype=np.complex128), (-5.0, 5.0, 10)),
        "crand": (rng.normal(size=50) + 1j * rng.normal(size=50), (-8.7, 9.1, 37)),
    }
    for cname, (c, (lo, hi, E)) in cases.items():
        cin = os.path.join(TMP, "c.bin")
        c.tofile(cin)
        gfa[f"{cname}_coef"] = c
        for t, tname in O.TYPE_NAMES.items():
            for lk in (0, 1):
                out = os.path.join(TMP, "g.bin")
                O.run_ref(mode="gfcoef", coefin=cin, scale=10.0,
                          window=f"{lo},{hi},{E}", type=tname, lookup=lk, out=out)
                gfa[f"{cname}_{tname}_lookup{lk}"] = np.fromfile(out, np.complex128)
    save("gf_types", dict(scale=10.0, windows={k: list(v[1]) for k, v in cases.items()}),
         **gfa)

    # Disordered 48x48 square lattice with spin (config 2 physics, small)
    m2 = L.square_spin(48, 48, t=1.0, J=0.5, disorder=1.0, seed=2)
    moments_fixture("square48_spin_disorder", m2, 12.0, 600, (24, 24, 0),
                    [(24, 24, 0), (24, 24, 1), (25, 24, 0)], keep_csr=False)

    # Nambu s-wave superconductor 12x12 (config 3 physics, small)
    m3 = L.nambu_swave(12, 12, t=1.0, mu=-1.0, delta=0.3, J=0.5, disorder=0.5, seed=3)
    s3 = round(m3.gershgorin_bound() * 1.05, 6)
    moments_fixture("nambu12", m3, s3, 500, (6, 6, 0),
                    [(6, 6, 0), (6, 6, 3), (7, 6, 0)])
    moments_fixture("nambu12_b", m3, s3, 500, (3, 9, 2), [(3, 9, 2)], keep_csr=False)

    # Honeycomb ribbon (config 4 family)
    m4 = L.honeycomb_ribbon(6, 9, t=2.7, disorder=0.5, seed=4)
    s4 = round(m4.gershgorin_bound() * 1.05, 6)
    moments_fixture("honeycomb6x9", m4, s4, 300, (3, 4, 1), [(3, 4, 1), (3, 4, 2)])

    # Cubic lattice (config 5 family)
    m5 = L.cubic(8, 8, 8, t=1.0, disorder=1.0, seed=5)
    moments_fixture("cubic8", m5, 7.0, 300, (4, 4, 4), [(4, 4, 4)], keep_csr=False)

    # LDOS sum rule on the ring of Test/.../PropertyExtractor/ChebyshevExpander.h:19-45
    mr = L.ring(10)
    pr, _ = ref_csr(mr, "ring10")
    ring = {}
    for frac in (0.9, 0.7):
        out = os.path.join(TMP, "ldos.bin")
        O.run_ref(mode="ldos", model=pr, scale=10.0, M=1000, broadening=1e-6,
                  window=f"{-10 * frac},{10 * frac},1000", lookup=0, patterns="0", out=out)
        ring[f"ldos_{frac:g}"] = np.fromfile(out, np.float64)
    save("ring10_ldos", dict(scale=10.0, M=1000, E=1000, fractions=[0.9, 0.7]), **ring)

    # sampleDOS on the 5x5 periodic lattice (site sequence from srand/rand)
    pp, _ = ref_csr(L.periodic_square(5), "p5s")
    out = os.path.join(TMP, "dos.bin")
    O.run_ref(mode="sampledos", model=pp, scale=10.0, M=200, broadening=1e-6,
              window="-9,9,200", lookup=1, n=7, seed=3, out=out)
    save("periodic5_sampledos", dict(scale=10.0, M=200, window=[-9, 9, 200], n=7,
                                     seed=3, broadening=1e-6),
         dos=np.fromfile(out, np.float64))

    if not only or "json" in only:
        json_fixture(p1, pp)
    if not only or "callers" in only:
        callers_fixture()


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
