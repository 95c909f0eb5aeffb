"""Expression-matrix text import/export (ImportExpressionMatrix::process,
src/core/importexpressionmatrix.cpp:20-170; ExportExpressionMatrix::process,
src/core/exportexpressionmatrix.cpp:20-99) against an independent Python
reading of the same rules: whitespace-split words, a header of sample names
(or numbered samples), the no-sample token as NaN, every other word read as
the nearest double (Python's float(), correctly rounded), transformed in
double and stored as float32."""
import math
import os
import subprocess

import numpy as np
import pytest

from kinc_amd import formats
from kinc_amd._native import KincError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def expect_value(word, nan, transform="none"):
    if nan and word == nan:
        return np.float32(np.nan)
    v = float(word)
    if transform == "natural logarithm":
        v = math.log(v) if v > 0 else (float("-inf") if v == 0 else float("nan"))
    elif transform == "logarithm base 2":
        v = math.log2(v) if v > 0 else (float("-inf") if v == 0 else float("nan"))
    elif transform == "logarithm base 10":
        v = math.log10(v) if v > 0 else (float("-inf") if v == 0 else float("nan"))
    with np.errstate(over="ignore"):
        return np.float32(v)


def expect_matrix(text, nan=None, samples=0, transform="none"):
    lines = text.split("\n")
    if text.endswith("\n"):
        lines = lines[:-1]
    names = []
    if samples == 0:
        while lines and samples == 0:
            names = lines.pop(0).split()
            samples = len(names)
    else:
        names = [str(i) for i in range(samples)]
    genes, rows = [], []
    for ln in lines:
        w = ln.split()
        assert len(w) == samples + 1
        genes.append(w[0])
        rows.append([expect_value(x, nan, transform) for x in w[1:]])
    return genes, names, np.array(rows, np.float32).reshape(len(rows), samples)


def same_bits(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return a.shape == b.shape and bool(np.all((np.isnan(a) & np.isnan(b)) |
                                              (a.view(np.uint32) == b.view(np.uint32))))


def roundtrip(tmp_path, text, **kw):
    src = tmp_path / "in.txt"
    src.write_bytes(text.encode())
    emx = str(tmp_path / "out.emx")
    G, S = formats.import_text(str(src), emx, **kw)
    data, meta = formats.read_emx(emx)
    return G, S, data, meta


def test_header_tokens_and_whitespace(tmp_path):
    text = ("s1  s2\ts3 \r\n"
            "geneA\t1.5\t2\tNA\r\n"
            "geneB   -0.25 1e3   +7\n"
            "geneC\t.5\t5.\t1E-3")          # last line without a newline
    G, S, data, meta = roundtrip(tmp_path, text, nan="NA")
    genes, names, want = expect_matrix(text.replace("\r", ""), nan="NA")
    assert (G, S) == (3, 3)
    assert meta["genes"] == ["geneA", "geneB", "geneC"] == genes
    assert meta["samples"] == ["s1", "s2", "s3"] == names
    assert meta["transform"] == "none"
    assert same_bits(data, want)
    assert np.isnan(data[0, 2])


def test_numbered_samples_without_header(tmp_path):
    text = "g0 1 2\ng1 3 4\n"
    G, S, data, meta = roundtrip(tmp_path, text, samples=2)
    assert (G, S) == (2, 2)
    assert meta["samples"] == ["0", "1"]
    assert same_bits(data, [[1, 2], [3, 4]])


@pytest.mark.parametrize("transform", formats.TRANSFORMS)
def test_transforms(tmp_path, transform):
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.uniform(0.01, 5000, 60), [1.0, 2.0, 10.0, 0.0]])
    words = ["%.9g" % v for v in vals]
    text = "s%s\n" % " s".join(str(i) for i in range(len(words))) + "g " + " ".join(words) + "\n"
    G, S, data, meta = roundtrip(tmp_path, text, transform=transform)
    assert meta["transform"] == transform
    want = np.array([expect_value(w, None, transform) for w in words], np.float32)
    assert same_bits(data[0], want)


def test_number_forms_bits(tmp_path):
    words = ["0", "-0", "0.1", "3.14159265358979323846264338", "123456789012345678901234",
             "1e22", "1e23", "9007199254740993", "4.9e-324", "1e-45", "7e-46", "3.4028235e38",
             "3.4028236e38", "1e39", "-1e-30", "0.000000000000000000000123",
             "16777217", "16777219", "0.30000000000000004", "nan", "inf", "+inf", "-inf",
             "00012.500", "1.17549435e-38", "2.5e-1", "1E+2"]
    text = " ".join("c%d" % i for i in range(len(words))) + "\n"
    text += "g " + " ".join(words) + "\n"
    G, S, data, meta = roundtrip(tmp_path, text)
    assert (G, S) == (1, len(words))
    want = np.array([expect_value(w, None) for w in words], np.float32)
    assert same_bits(data[0], want)


@pytest.mark.parametrize("bad", ["abc", "0x10", "1,5", "1e", "--1", "1.2.3", "e5", ".", "+"])
def test_unparsable_value_fails_with_reference_message(tmp_path, bad):
    text = "a b\ng1 1 %s\n" % bad
    with pytest.raises(KincError) as e:
        roundtrip(tmp_path, text)
    assert 'Failed to read expression value "%s" for gene g1.' % bad in str(e.value)


def test_wrong_field_count_fails_with_reference_message(tmp_path):
    text = "a b c\ng1 1 2 3\ng2 1 2\n"
    with pytest.raises(KincError) as e:
        roundtrip(tmp_path, text)
    assert ("incorrect amount of fields. Read in 2 fields when it should have been 3. "
            "Gene name is g2.") in str(e.value)
    # a blank line inside the data is a line with no fields
    with pytest.raises(KincError):
        roundtrip(tmp_path, "a b\ng1 1 2\n\ng2 3 4\n")


def test_first_bad_line_in_file_order_is_reported(tmp_path):
    # many lines so the parse runs on several threads; two bad lines
    rows = ["g%d %d %d" % (i, i, i + 1) for i in range(4000)]
    rows[3900] = "late 1"
    rows[1200] = "early 1 x"
    text = "a b\n" + "\n".join(rows) + "\n"
    with pytest.raises(KincError) as e:
        roundtrip(tmp_path, text, threads=8)
    assert 'Failed to read expression value "x" for gene early.' in str(e.value)


def test_threaded_import_matches_python_reading(tmp_path):
    rng = np.random.default_rng(11)
    G, S = 3000, 37
    x = rng.normal(8, 2, (G, S))
    x[rng.random((G, S)) < 0.1] = np.nan
    fmts = ["%.3f", "%.9g", "%.17g", "%e"]
    lines = ["\t".join("s%d" % j for j in range(S))]
    for g in range(G):
        f = fmts[g % 4]
        lines.append("gene%d\t" % g + "\t".join("NA" if np.isnan(v) else f % v for v in x[g]))
    text = "\n".join(lines) + "\n"
    for threads in (1, 8):
        Gi, Si, data, meta = roundtrip(tmp_path, text, nan="NA", threads=threads)
        assert (Gi, Si) == (G, S)
        _, _, want = expect_matrix(text, nan="NA")
        assert same_bits(data, want)


def test_export_format_and_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    data = rng.normal(0, 100, (20, 9)).astype(np.float32)
    data[2, 3] = np.nan
    data[4, 0] = 1e-7
    data[5, 1] = 123456789.0
    emx = str(tmp_path / "a.emx")
    genes = ["g%d" % i for i in range(20)]
    samples = ["s%d" % i for i in range(9)]
    formats.write_emx(emx, data, genes=genes, samples=samples)
    txt = str(tmp_path / "a.txt")
    formats.export_text(emx, txt, nan="NA")
    got = open(txt).read()
    want = "".join(s + "\t" for s in samples) + "\n"
    for g in range(20):
        want += genes[g] + "".join("\tNA" if np.isnan(v) else "\t%.12g" % float(v) for v in data[g])
        want += "\n"
    assert got == want
    # 12 significant digits carry every float32 exactly back
    emx2 = str(tmp_path / "b.emx")
    formats.import_text(txt, emx2, nan="NA")
    back, meta = formats.read_emx(emx2)
    assert same_bits(back, data)
    assert meta["genes"] == genes and meta["samples"] == samples


def test_export_undoes_log_transform(tmp_path):
    vals = [0.5, 1.0, 3.0, 1000.0, 12345.678]
    text = "a b c d e\ng " + " ".join(repr(v) for v in vals) + "\n"
    src = tmp_path / "in.txt"
    src.write_text(text)
    emx = str(tmp_path / "l.emx")
    formats.import_text(str(src), emx, transform="logarithm base 2")
    stored, _ = formats.read_emx(emx)
    txt = str(tmp_path / "l.txt")
    formats.export_text(emx, txt)
    out = open(txt).read().split("\n")[1].split("\t")[1:]
    want = ["%.12g" % float(np.float32(2.0 ** float(v))) for v in stored[0]]
    assert out == want
    assert np.allclose([float(v) for v in out], vals, rtol=1e-6)


def test_cli_import_export(tmp_path):
    exe = os.path.join(ROOT, "kinc_amd", "kinc")
    if not os.path.exists(exe):
        pytest.skip("kinc CLI not built")
    src = tmp_path / "c.txt"
    src.write_text("s1 s2\ngA 1 NA\ngB 2.5 -3\n")
    emx = str(tmp_path / "c.emx")
    r = subprocess.run([exe, "run", "import-emx", "--input", str(src), "--output", emx,
                        "--nan", "NA"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    data, meta = formats.read_emx(emx)
    assert same_bits(data, [[1, np.nan], [2.5, -3]])
    out = str(tmp_path / "c2.txt")
    r = subprocess.run([exe, "run", "export-emx", "--input", emx, "--output", out, "--nan", "NA"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert open(out).read() == "s1\ts2\t\ngA\t1\tNA\ngB\t2.5\t-3\n"
