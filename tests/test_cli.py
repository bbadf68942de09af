"""The drop-in command line: `kuramoto` / `kuramoto_64` (product, GPU) against
the oracle's restatement of the reference main() (CPU), same arguments, same
FILEBASE files (README.md:42-58; kuramoto.cu:164-265, :395-678).

File layouts checked (s = 4 or 8 bytes):
  frequencies.dat N s | adj.dat N^2 s | order.dat (1 + n_eval) s |
  thetas.dat (1 + n_eval) N s | couplings.dat (1 + n_eval) 2N s |
  times.dat accepted s | fs.dat (N + 2 + 2N) s
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC = os.path.join(ROOT, "oracle", "_build")
BIN = os.path.join(ROOT, "kuramoto_amd", "bin")


def run(prog, args, base, timeout=300):
    cmd = [prog] + args + [base]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, (cmd, out.stdout[-2000:], out.stderr[-2000:])
    return out.stdout


def load(base, suffix, dt):
    return np.fromfile(base + suffix, dtype=dt)


def check_layout(base, N, n_out, dt, dense=3, A=False):
    s = np.dtype(dt).itemsize
    assert os.path.getsize(base + "frequencies.dat") == N * s
    assert os.path.getsize(base + "order.dat") == n_out * s
    assert os.path.getsize(base + "fs.dat") == (3 * N + 2) * s
    if dense >= 1:
        assert os.path.getsize(base + "thetas.dat") == n_out * N * s
        assert os.path.getsize(base + "times.dat") % s == 0
    if dense >= 2:
        assert os.path.getsize(base + "couplings.dat") == n_out * 2 * N * s
    if dense >= 3 and not A:
        assert os.path.getsize(base + "adj.dat") == N * N * s
    header = open(base + ".out").readline().split()
    # dmd.py:40 reads N K t1 dt c seed from the first line
    assert len(header) == 6 and int(header[0]) == N


ARGS = ["-N", "120", "-K", "5", "-c", "3", "-t", "2", "-d", "0.25", "-s", "3", "-I", "6.283185307"]
N_OUT = 1 + 8


@pytest.mark.parametrize("prog,dt", [("kuramoto_oracle", np.float32), ("kuramoto_oracle_64", np.float64)])
def test_oracle_cli_layout(tmp_path, orc, prog, dt):
    base = str(tmp_path / "run")
    run(os.path.join(ORC, prog), ARGS + ["-n"], base)
    check_layout(base, 120, N_OUT, dt)
    M = load(base, "adj.dat", dt).reshape(120, 120)
    assert np.array_equal(M, M.T)


def test_oracle_cli_kill_resume(tmp_path, orc):
    """Kill after 7 accepted steps (--max-steps leaves the files as a killed
    process would), rerun with -R: order/thetas/times/couplings and the
    final fs.dat equal an uninterrupted run byte for byte (kuramoto.cu:471-494)."""
    prog = os.path.join(ORC, "kuramoto_oracle_64")
    a = str(tmp_path / "a")
    b = str(tmp_path / "b")
    run(prog, ARGS, a)
    run(prog, ARGS + ["--max-steps", "7"], b)
    assert os.path.getsize(b + "fs.dat") == (3 * 120 + 2) * 8
    run(prog, ARGS + ["-R"], b)
    for suf in ("order.dat", "thetas.dat", "times.dat", "couplings.dat", "fs.dat", "adj.dat", "frequencies.dat"):
        assert open(a + suf, "rb").read() == open(b + suf, "rb").read(), suf


def check_exact_final_continuation(prog, tmp_path, dt):
    """--exact-final (the optional fix-up of the reference's final fs.dat,
    kuramoto.cu:630-647 with dp45.cu:158-161): the last file holds the
    integrator's state, so -R with a later -t continues the run exactly: the
    accepted-step times, the grid rows (order, thetas, couplings) and the
    final fs.dat equal one uninterrupted run's byte for byte.  Without the
    flag the final fs.dat holds a phase vector interpolated at t_eval[last]
    beside a later t."""
    a, b, q = str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "q")
    base = ["-N", "120", "-K", "5", "-c", "3", "-d", "0.25", "-s", "3", "-I", "6.283185307"]
    run(prog, base + ["-t", "2", "--exact-final"], a)
    run(prog, base + ["-t", "2"], q)
    s = np.dtype(dt).itemsize
    fa, fq = load(a, "fs.dat", dt), load(q, "fs.dat", dt)
    t_a = float(fa[120])
    assert 2.0 <= t_a < 2.25  # the continuation's grid starts where the first run's ended
    assert fa[120] == fq[120] and fa[121] == fq[121]  # same integrator t and h
    assert not np.array_equal(fa[:120], fq[:120])  # the quirk: another phase vector
    n_first = 1 + 8
    run(prog, base + ["-t", "3", "--exact-final", "-R"], a)
    run(prog, base + ["-t", "3", "--exact-final"], b)
    for suf in ("times.dat", "fs.dat", "order.dat", "thetas.dat", "couplings.dat"):
        assert open(a + suf, "rb").read() == open(b + suf, "rb").read(), suf
    assert os.path.getsize(a + "order.dat") == (n_first + 4) * s


def test_oracle_cli_exact_final_continues(tmp_path, orc):
    check_exact_final_continuation(os.path.join(ORC, "kuramoto_oracle_64"), tmp_path, np.float64)


# ------------------------------------------------------------- GPU: product
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_cli_matches_oracle_cli(tmp_path, km, orc, prec):
    dt = np.float32 if prec == "f32" else np.float64
    prog = os.path.join(BIN, "kuramoto" if prec == "f32" else "kuramoto_64")
    oprog = os.path.join(ORC, "kuramoto_oracle" if prec == "f32" else "kuramoto_oracle_64")
    g, o = str(tmp_path / "g"), str(tmp_path / "o")
    args = ARGS + ["-n", "-a", "1e-6" if prec == "f32" else "1e-9"]
    print(f"case cli {prec}", flush=True)
    run(prog, args, g)
    run(oprog, args, o)
    check_layout(g, 120, N_OUT, dt)
    for suf in ("frequencies.dat", "adj.dat"):
        assert open(g + suf, "rb").read() == open(o + suf, "rb").read(), suf
    rg, ro = load(g, "order.dat", dt), load(o, "order.dat", dt)
    # r(t0): GPU block reduction vs the oracle's sequential sum (last-bit)
    assert abs(rg[0] - ro[0]) < (1e-14 if prec == "f64" else 1e-6)
    assert np.max(np.abs(rg - ro)) < 1e-3
    tg, to = load(g, "times.dat", dt), load(o, "times.dat", dt)
    assert abs(len(tg) - len(to)) <= max(1, int(0.02 * len(to)))
    thg = load(g, "thetas.dat", dt).reshape(N_OUT, 120)
    tho = load(o, "thetas.dat", dt).reshape(N_OUT, 120)
    assert np.array_equal(thg[0], tho[0])
    # every output row to t = 2, relative to max(1, |theta|): 1e-9 (FP64) /
    # 1e-5 (FP32; both sides sum in double and round cos / sin from double)
    tol = 1e-9 if prec == "f64" else 1e-5
    assert np.max(np.abs(thg - tho) / np.maximum(1, np.abs(tho))) < tol
    fg, fo = load(g, "fs.dat", dt), load(o, "fs.dat", dt)
    assert np.max(np.abs(fg[:120] - fo[:120]) / np.maximum(1, np.abs(fo[:120]))) < tol
    # Q7 (dp45.cu:160, kuramoto.cu:632-647): the final restart file holds the
    # integrator's t (the last accepted time, past t1) and h, then couplings
    assert fg[120] == tg[-1] and fo[120] == to[-1] and fg[120] >= dt(2.0)
    if prec == "f64":
        assert abs(fg[120] - fo[120]) < 1e-9 and abs(fg[121] - fo[121]) < 1e-6 * abs(fo[121])
        assert np.max(np.abs(fg[122:] - fo[122:])) < 1e-9
    cg = load(g, "couplings.dat", dt).reshape(N_OUT, 240)
    co = load(o, "couplings.dat", dt).reshape(N_OUT, 240)
    assert np.max(np.abs(cg - co)) < (1e-9 if prec == "f64" else 1e-5)


@pytest.mark.gpu
def test_cli_kill_resume_bit_exact(tmp_path, km):
    prog = os.path.join(BIN, "kuramoto_64")
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    print("case cli resume", flush=True)
    run(prog, ARGS, a)
    run(prog, ARGS + ["--max-steps", "7"], b)
    run(prog, ARGS + ["-R"], b)
    for suf in ("order.dat", "thetas.dat", "times.dat", "couplings.dat", "fs.dat"):
        assert open(a + suf, "rb").read() == open(b + suf, "rb").read(), suf


@pytest.mark.gpu
def test_cli_lowrank_and_reload(tmp_path, km, orc):
    """-A (factored coupling, no adj.dat) against the oracle's -A; then -R
    with the oracle's frequency and adjacency files gives the oracle's
    trajectory (identical inputs through -R)."""
    g, o = str(tmp_path / "g"), str(tmp_path / "o")
    args = ARGS + ["-A", "-a", "1e-9", "-D", "2"]
    print("case cli -A", flush=True)
    run(os.path.join(BIN, "kuramoto_64"), args, g)
    run(os.path.join(ORC, "kuramoto_oracle_64"), args, o)
    assert not os.path.exists(g + "adj.dat")
    check_layout(g, 120, N_OUT, np.float64, dense=2, A=True)
    assert np.max(np.abs(load(g, "order.dat", np.float64) - load(o, "order.dat", np.float64))) < 1e-9
    # -R: copy the oracle's inputs (adjacency, frequencies, a restart file)
    r = str(tmp_path / "r")
    o2 = str(tmp_path / "o2")
    run(os.path.join(ORC, "kuramoto_oracle_64"), ARGS + ["-a", "1e-9", "-s", "8"], o2)
    for suf in ("adj.dat", "frequencies.dat"):
        shutil.copy(o2 + suf, r + suf)
    th0 = load(o2, "thetas.dat", np.float64)[:120]
    np.concatenate([th0, [0.0, 0.02]]).tofile(r + "fs.dat")
    run(os.path.join(BIN, "kuramoto_64"), ARGS + ["-a", "1e-9", "-R", "-s", "99"], r)
    ro = load(o2, "order.dat", np.float64)[1:]
    rr = load(r, "order.dat", np.float64)
    assert len(rr) == len(ro) and np.max(np.abs(rr - ro)) < 1e-3
    assert "Using adjacency from file" in open(r + ".out").read()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_cli_high_rank_volcano(tmp_path, km, orc, prec):
    """Rank K = 40 > 16: the programs draw the volcano signs on the device
    (packed, kmt_set_factor_sign); the oracle's programs draw them on the
    host.  Dense: adj.dat byte-identical to the oracle's; -A: every output
    row within 1e-5 (FP32) / 1e-9 (FP64) of the oracle's -A run."""
    prog = os.path.join(BIN, "kuramoto" if prec == "f32" else "kuramoto_64")
    oprog = os.path.join(ORC, "kuramoto_oracle" if prec == "f32" else "kuramoto_oracle_64")
    dt = np.float32 if prec == "f32" else np.float64
    args = ["-N", "600", "-K", "40", "-c", "1.75", "-t", "1", "-d", "0.125", "-s", "5", "-n",
            "-a", "1e-6" if prec == "f32" else "1e-9"]
    print(f"case cli high-rank volcano {prec}", flush=True)
    g, o = str(tmp_path / "g"), str(tmp_path / "o")
    run(prog, args, g)
    run(oprog, args, o)
    assert open(g + "adj.dat", "rb").read() == open(o + "adj.dat", "rb").read()
    ga, oa = str(tmp_path / "ga"), str(tmp_path / "oa")
    run(prog, args + ["-A", "-D", "2"], ga)
    run(oprog, args + ["-A", "-D", "2"], oa)
    check_layout(ga, 600, 9, dt, dense=2, A=True)
    thg = load(ga, "thetas.dat", dt).reshape(9, 600)
    tho = load(oa, "thetas.dat", dt).reshape(9, 600)
    tol = 1e-9 if prec == "f64" else 1e-5
    assert np.max(np.abs(thg - tho) / np.maximum(1, np.abs(tho))) < tol
    assert np.max(np.abs(load(ga, "order.dat", dt) - load(oa, "order.dat", dt))) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_cli_rebuild_equals_dense(tmp_path, km, orc, prec):
    """`-A --rebuild` (adjacency tiles rebuilt in registers from the factor,
    no adj.dat) gives the dense run's files bit for bit, and follows the
    oracle's rebuilt-row run (`--rebuild`) to the order-parameter
    tolerance."""
    prog = os.path.join(BIN, "kuramoto" if prec == "f32" else "kuramoto_64")
    oprog = os.path.join(ORC, "kuramoto_oracle" if prec == "f32" else "kuramoto_oracle_64")
    dt = np.float32 if prec == "f32" else np.float64
    g, d, o = str(tmp_path / "g"), str(tmp_path / "d"), str(tmp_path / "o")
    args = ARGS + ["-D", "2", "-a", "1e-6" if prec == "f32" else "1e-9"]
    print(f"case cli rebuild {prec}", flush=True)
    run(prog, args + ["-A", "--rebuild"], g)
    run(prog, args, d)
    run(oprog, args + ["--rebuild"], o)
    assert not os.path.exists(g + "adj.dat")
    check_layout(g, 120, N_OUT, dt, dense=2, A=True)
    for suf in ("order.dat", "thetas.dat", "couplings.dat", "times.dat", "fs.dat"):
        assert open(g + suf, "rb").read() == open(d + suf, "rb").read(), suf
    tol = 1e-6 if prec == "f32" else 1e-9
    assert np.max(np.abs(load(g, "order.dat", dt) - load(o, "order.dat", dt))) < tol


def _dmd_load(filebase0):
    """The reading logic of dmd.py's load_data (dmd.py:16-60), restated with
    numpy: for n = 0, 1, ... while FILEBASE<n>.out exists, parse the first
    .out line as float64 (N K t1 dt c seed), take length = int(t1/dt) rows of
    float64 thetas.dat (shape (length, N)) and the float64 frequencies.dat,
    and move to the frame rotating at the mean frequency."""
    out, n = [], 0
    while os.path.exists(f"{filebase0}{n}.out"):
        fb = f"{filebase0}{n}"
        N, K, t1, dt, c, seed = np.array(open(fb + ".out").readline().split(), dtype=np.float64)
        N, length = int(N), int(t1 / dt)
        omega = np.memmap(fb + "frequencies.dat", dtype=np.float64, mode="r+", shape=(N,))
        theta = np.memmap(fb + "thetas.dat", dtype=np.float64, mode="r+", shape=(length, N))
        out.append((np.array(theta) - np.mean(omega) * dt * np.arange(length)[:, None], (N, K, t1, dt, c, seed)))
        n += 1
    return out


@pytest.mark.gpu
def test_dmd_load_data_reads_fp64_outputs(tmp_path, km, orc):
    """Downstream format consumer: two kuramoto_64 trajectories (FILEBASE0,
    FILEBASE1) read through dmd.py's load_data logic give the header values,
    (t1/dt) x N float64 theta rows starting with the initial phases, and the
    same rotating-frame data as the oracle's files read the same way."""
    g0, o0 = str(tmp_path / "g"), str(tmp_path / "o")
    for n, seed in enumerate((3, 4)):
        args = ["-N", "96", "-K", "5", "-c", "2.5", "-t", "2", "-d", "0.25", "-s", str(seed), "-n", "-a", "1e-9",
                "-I", "6.283185307"]
        print(f"case dmd trajectory {n}", flush=True)
        run(os.path.join(BIN, "kuramoto_64"), args, f"{g0}{n}")
        run(os.path.join(ORC, "kuramoto_oracle_64"), args, f"{o0}{n}")
    gl, ol = _dmd_load(g0), _dmd_load(o0)
    assert len(gl) == len(ol) == 2
    for (tg, hg), (to, ho) in zip(gl, ol):
        assert hg == ho and hg[0] == 96 and tg.shape == (8, 96)
        assert np.max(np.abs(tg - to)) < 1e-8
    assert np.array_equal(gl[0][0][0], np.fromfile(f"{g0}0thetas.dat", dtype=np.float64)[:96])


def test_dmd_load_data_on_oracle_outputs(tmp_path, orc):
    """dmd.py's reading logic on the oracle CLI's FP64 files (CPU): header,
    (t1/dt) rows of N, row 0 = initial phases, rotating frame."""
    o0 = str(tmp_path / "o")
    args = ["-N", "64", "-K", "5", "-c", "2.5", "-t", "1", "-d", "0.125", "-s", "5", "-n", "-a", "1e-9",
            "-I", "6.283185307"]
    run(os.path.join(ORC, "kuramoto_oracle_64"), args, f"{o0}0")
    (theta, hdr), = _dmd_load(o0)
    assert hdr == (64.0, 5.0, 1.0, 0.125, 2.5, 5.0) and theta.shape == (8, 64)
    raw = np.fromfile(f"{o0}0thetas.dat", dtype=np.float64).reshape(-1, 64)
    om = np.fromfile(f"{o0}0frequencies.dat", dtype=np.float64)
    assert raw.shape[0] == 9 and np.array_equal(theta[0], raw[0])
    assert np.allclose(theta[3], raw[3] - om.mean() * 0.125 * 3, rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("prog_name,dtype_opt", [("kuramoto", "bf16"), ("kuramoto", "f16"), ("kuramoto_64", "bf16")])
def test_cli_16bit_adjacency(tmp_path, km, prog_name, dtype_opt):
    """`--adj-dtype=bf16|f16` (A stored in 16 bits, FP32 accumulation; with
    kuramoto_64 the state, stage combinations and error norm stay FP64)
    writes the same files as the dense run and follows it: order.dat within
    1e-3, the same number of output rows, accepted steps within 2% (or 1)."""
    prog = os.path.join(BIN, prog_name)
    dt = np.float32 if prog_name == "kuramoto" else np.float64
    q, d = str(tmp_path / "q"), str(tmp_path / "d")
    args = ARGS + ["-D", "2", "-a", "1e-6" if prog_name == "kuramoto" else "1e-9"]
    print(f"case cli {prog_name} --adj-dtype={dtype_opt}", flush=True)
    run(prog, args + [f"--adj-dtype={dtype_opt}"], q)
    run(prog, args, d)
    check_layout(q, 120, N_OUT, dt, dense=2)
    assert np.max(np.abs(load(q, "order.dat", dt) - load(d, "order.dat", dt))) < 1e-3
    nq, nd = load(q, "times.dat", dt).size, load(d, "times.dat", dt).size
    assert abs(nq - nd) <= max(1, int(0.02 * nd)), (nq, nd)


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [["-F"], ["-f", "0.5"], ["-r", "1e-6"], ["--gauss"], ["-v"]],
                         ids=["fixed", "freq", "rtol", "gauss", "verbose"])
def test_cli_options_match_oracle_cli(tmp_path, km, orc, extra):
    """The remaining reference options (-F fixed step, -f frequency width,
    -r relative tolerance, -v verbose; kuramoto.cu:298-362) and --gauss on
    kuramoto_64 against the oracle's program with the same arguments: same
    frequencies, order.dat and phases within the FP64 tolerances, accepted
    steps within 2% (or 1).  --chunk (how many attempts run between host
    reads) must not change any file."""
    prog = os.path.join(BIN, "kuramoto_64")
    oprog = os.path.join(ORC, "kuramoto_oracle_64")
    dt = np.float64
    g, o, c = str(tmp_path / "g"), str(tmp_path / "o"), str(tmp_path / "c")
    args = ARGS + ["-D", "2", "-a", "1e-9"] + extra
    print(f"case cli option {' '.join(extra)}", flush=True)
    run(prog, args, g)
    run(oprog, args, o)
    run(prog, args + ["--chunk=3"], c)
    check_layout(g, 120, N_OUT, dt, dense=2)
    assert open(g + "frequencies.dat", "rb").read() == open(o + "frequencies.dat", "rb").read()
    assert np.max(np.abs(load(g, "order.dat", dt) - load(o, "order.dat", dt))) < 1e-9
    tg, to = load(g, "times.dat", dt), load(o, "times.dat", dt)
    assert abs(len(tg) - len(to)) <= max(1, int(0.02 * len(to)))
    thg = load(g, "thetas.dat", dt).reshape(N_OUT, 120)
    tho = load(o, "thetas.dat", dt).reshape(N_OUT, 120)
    assert np.max(np.abs(thg - tho) / np.maximum(1, np.abs(tho))) < 1e-9
    for suf in ("order.dat", "thetas.dat", "couplings.dat", "times.dat", "fs.dat"):
        assert open(g + suf, "rb").read() == open(c + suf, "rb").read(), suf


@pytest.mark.parametrize("args", [[], ["-h"], ["-N", "64", "-h"], ["BASE", "-h"]],
                         ids=["no-filebase", "h", "h-after-option", "h-after-filebase"])
@pytest.mark.parametrize("prog", ["kuramoto", "kuramoto_64"])
def test_cli_usage_exits_without_running(tmp_path, orc, prog, args):
    """kuramoto.cu:298-388: with -h, or without a FILEBASE, or with -h after
    the FILEBASE, main prints the usage text and exits 0 before any file or
    device is touched.  Runs on the CPU: no HIP call is made.  The text
    equals the oracle program's line for line; each adds one last line
    naming its own long-option extensions."""
    args = [str(tmp_path / "run") if a == "BASE" else a for a in args]
    got = subprocess.run([os.path.join(BIN, prog)] + args, capture_output=True, text=True, timeout=60)
    ref = subprocess.run([os.path.join(ORC, "kuramoto_oracle")] + args, capture_output=True, text=True, timeout=60)
    assert got.returncode == 0 and ref.returncode == 0
    lines, ref_lines = got.stdout.splitlines(), ref.stdout.splitlines()
    assert lines[0] == "usage:\tkuramoto [-hvnRFA] [-N N] [-K K] [-D D]"
    assert lines[-2] == ref_lines[-2] == "filebase is base file name for output. "
    assert lines[:-1] == ref_lines[:-1] and len(lines) == 24
    assert lines[-1] == ("extensions: --gauss --adj-dtype=native|bf16|f16|sym|sym-bf16|sym-f16 --rebuild "
                         "--max-steps=n --chunk=n --exact-final")
    assert os.listdir(tmp_path) == []


@pytest.mark.gpu
@pytest.mark.parametrize("prog,N", [("kuramoto", 300000), ("kuramoto_64", 200000)])
def test_cli_memory_check_stops_too_large_run(tmp_path, km, prog, N):
    """kuramoto.cu:424-438: when the device's free memory is below the
    estimate (here the N^2 adjacency alone, 360 / 320 GB, exceeds the 288 GB
    of HBM), main writes the header, the command line and the two memory
    lines to FILEBASE.out and returns 0 before allocating anything."""
    base = str(tmp_path / "big")
    print(f"case {prog} N={N} memory check", flush=True)
    out = run(os.path.join(BIN, prog), ["-N", str(N), "-K", "5", "-D", "0"], base, timeout=120)
    assert "GPU Memory low!" in out.splitlines()
    lines = open(base + ".out").read().splitlines()
    assert lines[0].split()[0] == str(N) and lines[-1] == "GPU Memory low!"
    fr, total, req = (int(v) for v in lines[-2].split()[2:])
    assert lines[-2].startswith("GPU Memory: ") and fr < req and total < req
    s = 4 if prog == "kuramoto" else 8
    assert req == (100 + 2 * 5) * N * s + N * N * s
    assert sorted(os.listdir(tmp_path)) == ["big.out"]


@pytest.mark.gpu
@pytest.mark.parametrize("prog_name,sym_opt,dense_opt", [("kuramoto", "sym", "native"), ("kuramoto_64", "sym", "native"),
                                                         ("kuramoto_64", "sym-bf16", "bf16")])
def test_cli_symmetric_storage(tmp_path, km, prog_name, sym_opt, dense_opt):
    """`--adj-dtype=sym` / `sym-bf16` (the upper block triangle only) runs
    the same method on the same matrix as `native` / `bf16`: adj.dat
    written from the stored triangle is byte-identical to the dense run's,
    and the run follows the dense one (only the order of the additions
    differs): FP64 order.dat and phases within 1e-9 (1e-5 with 16-bit
    storage, whose sums are FP32), FP32 order.dat within 1e-4, accepted
    steps within 2% (or 1).  A symmetric adj.dat reloads into symmetric
    storage (-R)."""
    prog = os.path.join(BIN, prog_name)
    dt = np.float32 if prog_name == "kuramoto" else np.float64
    s, d = str(tmp_path / "s"), str(tmp_path / "d")
    args = ARGS + ["-D", "3", "-a", "1e-6" if prog_name == "kuramoto" else "1e-9"]
    print(f"case cli {prog_name} --adj-dtype={sym_opt}", flush=True)
    run(prog, args + [f"--adj-dtype={sym_opt}"], s)
    run(prog, args + [f"--adj-dtype={dense_opt}"], d)
    check_layout(s, 120, N_OUT, dt, dense=3)
    for f in ("adj.dat", "frequencies.dat"):
        assert open(s + f, "rb").read() == open(d + f, "rb").read(), f
    tol = 1e-4 if dt == np.float32 else 1e-5 if sym_opt != "sym" else 1e-9
    assert np.max(np.abs(load(s, "order.dat", dt) - load(d, "order.dat", dt))) < tol
    if dt == np.float64:
        assert np.max(np.abs(load(s, "thetas.dat", dt) - load(d, "thetas.dat", dt))) < tol
    ns, nd = load(s, "times.dat", dt).size, load(d, "times.dat", dt).size
    assert abs(ns - nd) <= max(1, int(0.02 * nd)), (ns, nd)
    # -R: the dense run's adjacency and frequencies, its initial phases as
    # the restart file, into symmetric storage
    r = str(tmp_path / "r")
    for f in ("adj.dat", "frequencies.dat"):
        shutil.copy(d + f, r + f)
    np.concatenate([load(d, "thetas.dat", dt)[:120], np.array([0.0, 0.02], dt)]).astype(dt).tofile(r + "fs.dat")
    run(prog, args + ["-R", f"--adj-dtype={sym_opt}"], r)
    rr, rd = load(r, "order.dat", dt), load(d, "order.dat", dt)[1:]
    assert len(rr) == len(rd) and np.max(np.abs(rr - rd)) < 1e-3
    assert "Using adjacency from file" in open(r + ".out").read()


@pytest.mark.gpu
def test_cli_reference_quirks(tmp_path, km, orc):
    """Quirks of the reference command line that the drop-in keeps:
    Q3 (kuramoto.cu:294-339): "-nDv" reads "v" as the argument of -D, so the
    dense level is 0 and the run is not verbose: the same files as
    "-n -D 0", and the same set as the oracle program writes for "-nDv".
    Q4 (kuramoto.cu:498-500): the N initial-phase draws are made even with
    -I 0, so frequencies.dat is the same as with -I 6.28.
    Q15 (kuramoto.cu:317): -g is parsed with atof, "-g 0.7" is device 0.
    Q2: the default rank K is 6."""
    prog = os.path.join(BIN, "kuramoto")
    common = ["-N", "120", "-K", "5", "-c", "3", "-t", "1", "-d", "0.25", "-s", "3"]
    p = {k: str(tmp_path / k) for k in ("q3", "d0", "oq3", "i0", "i6", "g07", "g0")}
    print("case cli quirks Q3 Q4 Q15", flush=True)

    def files(base):
        d, stem = os.path.split(base)
        return sorted(f[len(stem):] for f in os.listdir(d) if f.startswith(stem))

    out_q3 = run(prog, common + ["-nDv"], p["q3"])
    out_d0 = run(prog, common + ["-n", "-D", "0"], p["d0"])
    run(os.path.join(ORC, "kuramoto_oracle"), common + ["-nDv"], p["oq3"])
    assert files(p["q3"]) == files(p["d0"]) == files(p["oq3"])
    assert "thetas.dat" not in files(p["q3"]) and len(out_q3.splitlines()) == len(out_d0.splitlines())
    assert np.array_equal(load(p["q3"], "order.dat", np.float32), load(p["d0"], "order.dat", np.float32))
    run(prog, common + ["-I", "0", "-D", "0"], p["i0"])
    run(prog, common + ["-I", "6.283185307", "-D", "0"], p["i6"])
    assert open(p["i0"] + "frequencies.dat", "rb").read() == open(p["i6"] + "frequencies.dat", "rb").read()
    run(prog, common + ["-g", "0.7", "-D", "0"], p["g07"])
    run(prog, common + ["-g", "0", "-D", "0"], p["g0"])
    assert open(p["g07"] + "order.dat", "rb").read() == open(p["g0"] + "order.dat", "rb").read()
    # Q2 (kuramoto.cu:278): without -K the rank is 6 (the usage text says 5)
    k6 = str(tmp_path / "k6")
    run(prog, [x for x in common if x not in ("-K", "5")] + ["-D", "0"], k6)
    assert open(k6 + ".out").readline().split()[:2] == ["120", "6"]


@pytest.mark.gpu
def test_cli_reference_quirks_restart_and_overshoot(tmp_path, km, orc):
    """Q5 (kuramoto.cu:475-489, 501-503): a FILEBASEfs.dat holding only N
    values (what the startup write leaves) is not a restart file, so -R
    falls back to the seed's random phases and the run equals the run
    without -R (adjacency and frequencies regenerated from the seed too,
    their files being absent).
    Q6 (dp45.cu:153-155): the last step is not clamped, so the last
    accepted time in times.dat reaches or passes t1, as in the oracle
    program's run."""
    prog = os.path.join(BIN, "kuramoto")
    common = ["-N", "120", "-K", "5", "-c", "3", "-t", "1", "-d", "0.25", "-s", "3", "-I", "6.283185307"]
    a, r, o = str(tmp_path / "a"), str(tmp_path / "r"), str(tmp_path / "o")
    print("case cli quirks Q5 Q6", flush=True)
    run(prog, common + ["-D", "1"], a)
    np.random.default_rng(0).standard_normal(120).astype(np.float32).tofile(r + "fs.dat")
    run(prog, common + ["-D", "1", "-R"], r)
    assert open(a + "order.dat", "rb").read() == open(r + "order.dat", "rb").read()
    assert open(a + "thetas.dat", "rb").read() == open(r + "thetas.dat", "rb").read()
    ta = load(a, "times.dat", np.float32)
    assert ta[-1] >= np.float32(1.0) and np.all(np.diff(ta) > 0)
    run(os.path.join(ORC, "kuramoto_oracle"), common + ["-D", "1"], o)
    to = load(o, "times.dat", np.float32)
    # FP32 step sequences of two summation orders drift apart slightly (the
    # CLI comparison above allows 2% in the times): both overshoot t1
    assert to[-1] >= np.float32(1.0) and abs(len(ta) - len(to)) <= 1
    assert abs(float(ta[-1]) - float(to[-1])) <= 0.02 * float(to[-1])


@pytest.mark.gpu
@pytest.mark.parametrize("prog_name,dt", [("kuramoto", np.float32), ("kuramoto_64", np.float64)])
def test_cli_exact_final_continues(tmp_path, km, prog_name, dt):
    print(f"case cli exact-final {prog_name}", flush=True)
    check_exact_final_continuation(os.path.join(BIN, prog_name), tmp_path, dt)
