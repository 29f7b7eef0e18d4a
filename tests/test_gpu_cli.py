"""The mrf command-line tool end to end on the GPU (src/mrf.cu:34-386 drop-in):
CSV sequence + RA image + RA B1 map in, RA atoms / params / maps out, each
checked against the library entry points and the FP64 oracle."""
import subprocess

import numpy as np
import pytest

import oracle
import snapmrf_amd as sm
from snapmrf_amd import synth

pytestmark = pytest.mark.gpu

GRID = ("100:100:1500", "20:20:300", "-20:20:20", "0.9:0.1:1.1")
T, W = 300, 30


def _setup(gpu, tmp_path, nx=24, ny=20):
    raw = synth.csv_columns(synth.sequence(T, seed=5), "epg_ssfp")
    csv = tmp_path / "seq.csv"
    synth.write_csv(str(csv), raw)
    params, A, natoms = sm.build_params(*GRID)
    d_params = gpu.from_numpy(params).cuda()
    d_atoms = gpu.empty((A, T), dtype=gpu.complex64, device="cuda")
    # the CSV round trip (%.9g) is exact for float32, so these are the CLI's inputs
    sm.MRF_dict(d_atoms, d_params, raw, T, T, A, natoms, 4, W, "epg_ssfp")
    atoms = d_atoms.cpu().numpy()
    b1 = synth.b1_map(nx, ny, seed=3).astype(np.float32)          # [ny, nx]
    slice_b1 = params[::natoms, 3]
    ks = np.array([oracle.b1_slice(float(v), slice_b1) for v in b1.ravel()])
    rng = np.random.default_rng(4)
    V = nx * ny
    truth = ks * natoms + rng.integers(0, natoms, V)
    pd = rng.uniform(0.4, 1.0, V).astype(np.float32)
    sig = atoms[truth] * pd[:, None]
    sig = (sig + (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)) * 0.002).astype(np.complex64)
    img = tmp_path / "img.ra"
    sm.ra_write(str(img), sig.reshape(ny, nx, T))          # RA dims (T, nx, ny)
    b1f = tmp_path / "b1.ra"
    sm.ra_write(str(b1f), b1)                             # RA dims (nx, ny)
    return raw, params, A, natoms, atoms, sig, b1.ravel(), truth, csv, img, b1f


def _run(tmp_path, csv, img, b1f, *extra):
    out = {k: str(tmp_path / f"{k}.ra") for k in ("atoms", "maps", "params")}
    cmd = [sm.CLI_PATH, "-t", GRID[0], "-s", GRID[1], "-b", GRID[2], "-r", GRID[3], "-w", str(W),
           "-T", str(T), "-M", str(b1f), "-a", out["atoms"], "-m", out["maps"], "-p", out["params"],
           *extra, str(csv), str(img)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return {k: sm.ra_read(v) for k, v in out.items()}, r.stdout


def test_cli_end_to_end(gpu, tmp_path):
    raw, params, A, natoms, atoms, sig, b1, truth, csv, img, b1f = _setup(gpu, tmp_path)
    res, stdout = _run(tmp_path, csv, img, b1f)
    assert f"NATOMS: {A}" in stdout
    a_cli, ha = res["atoms"]
    assert ha["dims"] == [T, A] and np.array_equal(a_cli, atoms)
    p_cli, hp = res["params"]
    assert hp["dims"] == [4, A] and np.array_equal(p_cli, params)
    m_cli, hm = res["maps"]
    assert hm["dims"] == [5, len(sig)]
    # the library's FP32 matcher on the same inputs
    d_maps = gpu.empty((len(sig), 5), device="cuda")
    d_idx = gpu.empty(len(sig), dtype=gpu.int32, device="cuda")
    sm.match_ex(sm.MRF_MATCH_FP32, gpu.from_numpy(atoms).cuda(), gpu.from_numpy(params).cuda(), T,
                A, natoms, gpu.from_numpy(sig).cuda(), len(sig), d_maps,
                d_b1map=gpu.from_numpy(b1).cuda(), d_idx=d_idx)
    gpu.cuda.synchronize()
    assert np.array_equal(m_cli, d_maps.cpu().numpy())
    # and the FP64 oracle per B1 slice on clear voxels
    idx = d_idx.cpu().numpy()
    slice_b1 = params[::natoms, 3]
    ks = np.array([oracle.b1_slice(float(v), slice_b1) for v in b1])
    for k in np.unique(ks):
        sel = ks == k
        ridx, rpd, rsec = oracle.match_slice(atoms[k * natoms:(k + 1) * natoms], sig[sel])
        clear = (rpd - rsec) / rpd > 1e-5
        assert np.array_equal(idx[sel][clear], ridx[clear] + k * natoms)
    assert np.mean(idx == truth) > 0.95


@pytest.mark.parametrize("precision", ["split2", "bf16_top16", "bf16_top32", "bf16_top64"])
def test_cli_low_precision(gpu, tmp_path, precision):
    raw, params, A, natoms, atoms, sig, b1, truth, csv, img, b1f = _setup(gpu, tmp_path)
    res32, _ = _run(tmp_path, csv, img, b1f)
    res, stdout = _run(tmp_path, csv, img, b1f, "--precision", precision)
    m32, m = res32["maps"][0], res["maps"][0]
    same = np.all(m32[:, :4] == m[:, :4], axis=1)
    assert np.mean(same) >= 0.999


def test_cli_natoms_mode_without_b1_map(gpu, tmp_path):
    """-N grid (src/mrf.cu:179-203: log-spaced T1/T2, symmetric-log B0, 11 B1
    slices) and matching without a B1 map, where the slice with the largest
    PD wins (src/functions.cu:1386-1416), i.e. the first-index argmax over all
    atoms."""
    Tn, Wn, N = 60, 10, 3000
    raw = synth.csv_columns(synth.sequence(Tn, seed=6), "epg_ssfp")
    csv = tmp_path / "seq.csv"
    synth.write_csv(str(csv), raw)
    params, A, natoms = sm.params_from_axes(*sm.grid_natoms_mode(N))
    assert A // natoms == 11
    d_params = gpu.from_numpy(params).cuda()
    d_atoms = gpu.empty((A, Tn), dtype=gpu.complex64, device="cuda")
    sm.MRF_dict(d_atoms, d_params, raw, Tn, Tn, A, natoms, 4, Wn, "epg_ssfp")
    atoms = d_atoms.cpu().numpy()
    rng = np.random.default_rng(8)
    nx, ny = 12, 10
    V = nx * ny
    truth = rng.integers(0, A, V)
    sig = atoms[truth] * rng.uniform(0.4, 1.0, V).astype(np.float32)[:, None]
    sig = (sig + (rng.normal(size=sig.shape) + 1j * rng.normal(size=sig.shape)) * 0.002).astype(np.complex64)
    img = tmp_path / "img.ra"
    sm.ra_write(str(img), sig.reshape(ny, nx, Tn))
    out = {k: str(tmp_path / f"{k}.ra") for k in ("atoms", "maps", "params")}
    r = subprocess.run([sm.CLI_PATH, "-N", str(N), "-w", str(Wn), "-T", str(Tn),
                        "-a", out["atoms"], "-m", out["maps"], "-p", out["params"],
                        str(csv), str(img)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"NATOMS: {A}" in r.stdout
    assert np.array_equal(sm.ra_read(out["params"])[0], params)
    assert np.array_equal(sm.ra_read(out["atoms"])[0], atoms)
    m_cli = sm.ra_read(out["maps"])[0]
    d_maps = gpu.empty((V, 5), device="cuda")
    sm.MRF_match(d_maps, sig, d_atoms, d_params, Tn, Tn, A, natoms, V, 4, 5, None)
    assert np.array_equal(m_cli, d_maps.cpu().numpy())
    # FP64 oracle: global argmax, on voxels whose top-2 gap is clear
    ridx, rpd, rsec = oracle.match_slice(atoms, sig)
    clear = (rpd - rsec) / rpd > 1e-5
    assert np.all(m_cli[clear, :4] == params[ridx[clear]])
    assert np.allclose(m_cli[clear, 4], rpd[clear], rtol=1e-4)


def test_cli_stream_equals_whole_dictionary(gpu, tmp_path):
    """--stream generates and matches block by block: the same maps as the
    default run, without -a."""
    raw, params, A, natoms, atoms, sig, b1, truth, csv, img, b1f = _setup(gpu, tmp_path)
    ref, _ = _run(tmp_path, csv, img, b1f)
    out = {k: str(tmp_path / f"s_{k}.ra") for k in ("maps", "params")}
    r = subprocess.run([sm.CLI_PATH, "-t", GRID[0], "-s", GRID[1], "-b", GRID[2], "-r", GRID[3],
                        "-w", str(W), "-T", str(T), "-M", str(b1f), "-m", out["maps"],
                        "-p", out["params"], "--stream", str(csv), str(img)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(sm.ra_read(out["maps"])[0], ref["maps"][0])
    assert np.array_equal(sm.ra_read(out["params"])[0], params)
    # -a needs the whole dictionary
    bad = subprocess.run([sm.CLI_PATH, "--stream", "-a", str(tmp_path / "x.ra"), str(csv), str(img)],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0


def test_cli_svd(gpu, tmp_path):
    """--svd FRAC: per-B1-slice bases (the B1 map picks each voxel's basis),
    rank by the energy rule, compressed FP32 matching."""
    raw, params, A, natoms, atoms, sig, b1, truth, csv, img, b1f = _setup(gpu, tmp_path)
    ref, _ = _run(tmp_path, csv, img, b1f)
    res, stdout = _run(tmp_path, csv, img, b1f, "--svd", "0.99999")
    assert "SVD rank:" in stdout and "per B1 slice" in stdout
    m, m32 = res["maps"][0], ref["maps"][0]
    # every map row is a grid point and nearly all equal the uncompressed pick
    grid_rows = {tuple(r) for r in params.tolist()}
    assert all(tuple(r) in grid_rows for r in m[:, :4].tolist())
    assert np.mean(np.all(m[:, :4] == m32[:, :4], axis=1)) >= 0.95


def test_cli_stream_low_precision(gpu, tmp_path):
    """--stream with --precision bf16_top32: block-streamed BF16 matching;
    its maps equal the FP32 run's on nearly every voxel."""
    raw, params, A, natoms, atoms, sig, b1, truth, csv, img, b1f = _setup(gpu, tmp_path)
    ref, _ = _run(tmp_path, csv, img, b1f)
    out = str(tmp_path / "s_maps.ra")
    r = subprocess.run([sm.CLI_PATH, "-t", GRID[0], "-s", GRID[1], "-b", GRID[2], "-r", GRID[3],
                        "-w", str(W), "-T", str(T), "-M", str(b1f), "-m", out, "--stream",
                        "--precision", "bf16_top32", str(csv), str(img)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m, m32 = sm.ra_read(out)[0], ref["maps"][0]
    assert np.mean(np.all(m[:, :4] == m32[:, :4], axis=1)) >= 0.999
