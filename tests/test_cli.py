"""The standalone command line (cudaGSEA/src/cudaGSEA.cu, flags of
cudaGSEA/src/HELP.txt) and the Broad-format loaders behind it
(read_data_files.cuh:46-321).  Fixtures: the reference's own hallmark GMT
(data/Pathways/h.all.v5.0.symbols.gmt) and GSE19429 class file
(data/GSE19429/GSE19429_series.cls), copied to tests/golden/; the GCT the
reference used is not shipped, so the tests write seeded GCT tables over the
hallmark symbols.  CPU tests use -parse-only and the error paths; the GPU test
checks the RESULT lines against the library's Python mirror and the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cudagsea_amd", "cudagsea")
GOLD = os.path.join(ROOT, "tests", "golden")
GMT = os.path.join(GOLD, "h.all.v5.0.symbols.gmt")
CLS = os.path.join(GOLD, "GSE19429_series.cls")


def need_cli():
    if not os.path.exists(CLI):
        pytest.skip("command line not built (python -m cudagsea_amd.build)")


def run(*args):
    p = subprocess.run([CLI] + [str(a) for a in args], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    return p.returncode, p.stdout


def hallmark_symbols():
    syms = []
    for line in open(GMT):
        syms.extend(t.strip() for t in line.rstrip("\n").split("\t")[2:])
    return sorted(set(s for s in syms if s))


def write_gct(path, genes, S, seed=0, dup=False, header="#1.2"):
    rng = np.random.default_rng(seed)
    X = np.exp(5.0 + rng.standard_normal((len(genes), S)))
    with open(path, "w") as f:
        f.write(header + "\n%d\t%d\n" % (len(genes), S))
        f.write("NAME\tDescription\t" + "\t".join("S%03d" % i for i in range(S)) + "\n")
        for i, g in enumerate(genes):
            name = genes[0] if (dup and i == len(genes) - 1) else g
            f.write(name + "\tna\t" + "\t".join("%.6f" % v for v in X[i]) + "\n")
    return X


def test_parse_only_reference_fixtures(tmp_path):
    need_cli()
    genes = hallmark_symbols()
    gct = tmp_path / "e.gct"
    write_gct(gct, genes, 200)
    rc, out = run("-res", gct, "-cls", CLS, "-gmx", GMT, "-parse-only")
    assert rc == 0, out
    assert "(%d genes x 200 samples)" % len(genes) in out
    labels = open(CLS).read().split("\n")[2].split()
    nA = labels.count("0")
    assert "(%d / %d)" % (nA, 200 - nA) in out
    assert "(50 sets)" in out


@pytest.mark.parametrize("args,code", [
    ([], 1),
    (["-cls", CLS, "-gmx", GMT], 11),
    (["-res", "X.gct", "-gmx", GMT], 13),
    (["-res", "/nonexistent.gct", "-cls", CLS, "-gmx", GMT], 14),
])
def test_flag_errors(args, code):
    need_cli()
    rc, out = run(*args)
    assert rc == code, out


def test_value_errors(tmp_path):
    need_cli()
    gct = tmp_path / "e.gct"
    write_gct(gct, hallmark_symbols()[:300], 200)
    base = ["-res", gct, "-cls", CLS, "-gmx", GMT]
    assert run("-res", gct, "-cls", "/nonexistent.cls", "-gmx", GMT)[0] == 15
    assert run("-res", gct, "-cls", CLS, "-gmx", "/nonexistent.gmt")[0] == 16
    assert run(*base, "-nperm", "12x")[0] == 17
    assert run(*base, "-order", "sideways")[0] == 18
    assert run(*base, "-dump")[0] == 19
    assert run(*base, "-precision", "half")[0] == 20
    assert run(*base, "-gpu", "0", "-cpu")[0] == 22
    assert run(*base, "-cpu")[0] == 22


def test_loader_errors(tmp_path):
    need_cli()
    genes = hallmark_symbols()[:200]
    g1 = tmp_path / "dup.gct"
    write_gct(g1, genes, 200, dup=True)
    rc, out = run("-res", g1, "-cls", CLS, "-gmx", GMT, "-parse-only")
    assert rc == 3 and "duplicate" in out           # DUPLICATE_GENE_SYMBOL
    g2 = tmp_path / "hdr.gct"
    write_gct(g2, genes, 200, header="#1.3")
    assert run("-res", g2, "-cls", CLS, "-gmx", GMT, "-parse-only")[0] == 4
    g3 = tmp_path / "ok.gct"
    write_gct(g3, genes, 200)
    c1 = tmp_path / "three.cls"
    c1.write_text("200 3 1\n# A B C\n" + " ".join(["0"] * 200) + "\n")
    assert run("-res", g3, "-cls", c1, "-gmx", GMT, "-parse-only")[0] == 6
    c2 = tmp_path / "unknown.cls"
    c2.write_text("200 2 1\n# A B\n" + " ".join(["A"] * 199 + ["Q"]) + "\n")
    assert run("-res", g3, "-cls", c2, "-gmx", GMT, "-parse-only")[0] == 7
    c3 = tmp_path / "short.cls"
    c3.write_text("100 2 1\n# A B\n" + " ".join(["A", "B"] * 50) + "\n")
    assert run("-res", g3, "-cls", c3, "-gmx", GMT, "-parse-only")[0] == 9
    # symbolic labels map to 0/1 by their order in the header line
    c4 = tmp_path / "sym.cls"
    c4.write_text("200 2 1\n# CASE CTRL\n" + " ".join(["CASE"] * 60 + ["CTRL"] * 140) + "\n")
    rc, out = run("-res", g3, "-cls", c4, "-gmx", GMT, "-parse-only")
    assert rc == 0 and "(60 / 140)" in out


RESULT = re.compile(r"RESULT: ES: (\S+) NES: (\S+) NP: (\S+) FWER: (\S+)\t\((.*)\)")


@pytest.mark.gpu
def test_cli_results_match_library_and_oracle(gpu, tmp_path):
    """RESULT lines equal the library's GSEA() mirror (same engine) to the
    printed digits, and the oracle within the ES tolerance."""
    import oracle
    genes = hallmark_symbols()[:1500]
    S = 200
    gct = tmp_path / "e.gct"
    write_gct(gct, genes, S, seed=3)
    P = 256
    print("cli run", len(genes), S, P, flush=True)
    rc, out = run("-res", gct, "-cls", CLS, "-gmx", GMT, "-nperm", P,
                  "-metric", "stable_signal2noise")
    assert rc == 0, out[-2000:]
    rows = [RESULT.match(l) for l in out.splitlines() if l.startswith("RESULT:")]
    assert len(rows) == 50 and all(rows)
    got = np.array([[float(m.group(i)) for i in range(1, 5)] for m in rows])
    names = [m.group(5) for m in rows]
    # the Python mirror of the R entry point on the same parsed data
    import cudagsea_amd as cg
    X = np.loadtxt(gct, skiprows=3, usecols=range(2, 2 + S), dtype=np.float64)
    labels = np.array(open(CLS).read().split("\n")[2].split(), dtype=np.int64)
    sets = {}
    for line in open(GMT):
        t = line.rstrip("\n").split("\t")
        sets[t[0]] = [x.strip() for x in t[2:] if x.strip()]
    import pandas as pd
    Xf = X.astype(np.float32)   # the CLI casts the GCT's values to float
    df = cg.GSEA(pd.DataFrame(Xf, index=genes), labels, sets, P, "stable_signal2noise")
    want = df.loc[names, ["ES", "NES", "NP", "FWER"]].to_numpy()
    # printed with 6 decimals
    np.testing.assert_allclose(got, want, atol=1.5e-6)
    # oracle on the mapped CSR
    sym = {g: i for i, g in enumerate(genes)}
    off, gl = [0], []
    for n in names:
        ids = sorted({sym[s] for s in sets[n] if s in sym})
        gl.extend(ids)
        off.append(len(gl))
    st, _ = oracle.gsea(np.ascontiguousarray(Xf.T), labels.astype(np.uint8),
                        np.array(off, np.int32), np.array(gl, np.int32), P,
                        "stable_signal2noise")
    np.testing.assert_allclose(got[:, 0], st[0], atol=2e-6)
    np.testing.assert_allclose(got[:, 1], st[1], rtol=1e-5, atol=2e-6)


@pytest.mark.gpu
def test_cli_ascending_double_dump(gpu, tmp_path):
    """-order ascending, -precision double, -dump and -gpu 0: RESULT lines
    equal the library's fp64 ascending run on the same parsed data, the dump
    file has the reference's name and holds that run's ES matrix, and the
    observed column matches the oracle."""
    import oracle
    import parity
    import cudagsea_amd as cg
    need_cli()
    genes = hallmark_symbols()[:1500]
    S = 200
    gct = tmp_path / "e.gct"
    write_gct(gct, genes, S, seed=5)
    P = 128
    prefix = tmp_path / "run"
    print("cli ascending double dump", len(genes), S, P, flush=True)
    rc, out = run("-res", gct, "-cls", CLS, "-gmx", GMT, "-nperm", P,
                  "-metric", "twopass_t_test", "-order", "ascending",
                  "-precision", "double", "-dump", prefix, "-gpu", 0)
    assert rc == 0, out[-2000:]
    rows = [RESULT.match(l) for l in out.splitlines() if l.startswith("RESULT:")]
    assert len(rows) == 50 and all(rows)
    got = np.array([[float(m.group(i)) for i in range(1, 5)] for m in rows])
    names = [m.group(5) for m in rows]
    X = np.loadtxt(gct, skiprows=3, usecols=range(2, 2 + S), dtype=np.float64)
    labels = np.array(open(CLS).read().split("\n")[2].split(), dtype=np.uint8)
    sets = {}
    for line in open(GMT):
        t = line.rstrip("\n").split("\t")
        sets[t[0]] = [x.strip() for x in t[2:] if x.strip()]
    sym = {g: i for i, g in enumerate(genes)}
    off, gl = [0], []
    for n in names:
        gl.extend(sorted({sym[s] for s in sets[n] if s in sym}))
        off.append(len(gl))
    off, gl = np.array(off, np.int32), np.array(gl, np.int32)
    Xs = np.ascontiguousarray(X.T)            # sample-major
    st, es = cg.run(Xs, labels, off, gl, P, "twopass_t_test", ascending=True,
                    double=True, want_es=True)
    np.testing.assert_allclose(got, st[:4].T, atol=1.5e-6)   # printed with 6 decimals
    dump = "%s_%d_%d_64.es" % (prefix, len(names), P)
    assert os.path.exists(dump), os.listdir(tmp_path)
    d = cg.read_dump(dump)
    assert d.dtype == np.float64 and d.shape == (len(names), P)
    assert parity.bits_equal(d, es).all()
    r = oracle.correlate(Xs, labels, P, "twopass_t_test", sample_major=True, double=True)
    sk, idx = oracle.pathway_sort(r, ascending=True)
    want, hi, lo = oracle.scores_ext(sk, idx, off, gl)
    parity.assert_es(d, want, hi, lo)
