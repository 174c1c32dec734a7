"""cook's playback file (.plb) written by cookio.PlaybackWriter, pinned by
the file cook wrote in tests/golden/spce216_plb (216 SPC/E, 4 SHAKE steps,
-y216, dt.plb=h): frames 0-3 are the float32 positions of the force calls
dumped by -v128 (dumps 0, 2, 3, 4), so the writer fed those positions
gives cook's bytes exactly; the last frame, after the last step, comes
from one oracle step (positions to 1e-10 of cook's, so within one float
ulp)."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, cookio

FRAME_DUMPS = [0, 2, 3, 4]


@pytest.fixture(scope="module")
def case():
    return oracle.load_fixture("spce216_plb")


def test_writer_gives_cooks_bytes(case, tmp_path):
    d, meta = case
    ns = d["r"].shape[1]
    raw = d["plb"].tobytes()
    path = tmp_path / "md.plb"
    with cookio.PlaybackWriter(path, ns) as w:
        for k in FRAME_DUMPS:
            w.write(d["L"], d["r"][k])
    got = path.read_bytes()
    assert len(got) == 8 + 4 * 12 * (ns + 1)
    assert got == raw[:len(got)]
    # the last frame: r(t_4) from the oracle's step on r(t_3), r1 = r(t_3) - r(t_2)
    bs = oracle.fixture_system(dict(meta, epsr=0.05, epsk=0.05))
    r = d["r"][4].copy()
    r1 = d["r"][4] - d["r"][3]
    p = abi.Step(h=meta["h"], eps=1e-12, maxit=0, thermostat=0, T=298, tau=0.1, nf=1,
                 measure_every=1)
    oracle.md_step(bs, r, r1, p)
    with cookio.PlaybackWriter(path, ns, append=True) as w:
        w.write(d["L"], r)
    ns_, flag, L, frames = cookio.read_plb(path)
    ref = np.frombuffer(raw[8:], "<f4").reshape(-1, ns + 1, 3)
    assert frames.shape == (5, ns, 3) and flag == -3 and ns_ == ns
    last = frames[4]
    ulp = np.spacing(np.abs(ref[4, 1:]).astype(np.float32))
    assert np.all(np.abs(last - ref[4, 1:]) <= ulp)


def test_reader_and_variants(case, tmp_path):
    d, meta = case
    ns = d["r"].shape[1]
    path = tmp_path / "c.plb"
    raw = d["plb"].tobytes()
    (tmp_path / "cook.plb").write_bytes(raw)
    ns_, flag, L, frames = cookio.read_plb(tmp_path / "cook.plb")
    assert (ns_, flag) == (ns, -3.0) and frames.shape == (5, ns, 3)
    np.testing.assert_array_equal(L, np.float32(meta["L"]))
    np.testing.assert_array_equal(frames[0], d["r"][0].astype(np.float32))
    # centred file (cook's -n4): flag -6, sites shifted by -L/2
    with cookio.PlaybackWriter(path, ns, centered=True) as w:
        w.write(d["L"], d["r"][0])
    _, flag, _, fr = cookio.read_plb(path)
    assert flag == -6.0
    np.testing.assert_array_equal(fr[0], (d["r"][0] - np.asarray(d["L"]) / 2).astype(np.float32))
    with pytest.raises(cookio.CookInputError):
        cookio.PlaybackWriter(path, ns + 3, append=True)
    (tmp_path / "bad.plb").write_bytes(raw[:-4])
    with pytest.raises(cookio.CookInputError):
        cookio.read_plb(tmp_path / "bad.plb")


@pytest.mark.gpu
def test_engine_trajectory_writes_cooks_playback(case, tmp_path):
    """the engine's FP64 SHAKE steps from cook's state after step 1,
    downloaded and written each step, give cook's frames 1-4 within one
    float ulp"""
    from macsimus_amd.engine import Engine
    from macsimus_amd.system import engine_provider
    d, meta = case
    ns = d["r"].shape[1]
    bs = oracle.fixture_system(dict(meta, epsr=0.05, epsk=0.05), provider_=engine_provider())
    e = Engine(bs, abi.MDX_FP64)
    e.upload_state(d["r"][2], d["r"][2] - d["r"][1])
    p = abi.Step(h=meta["h"], eps=1e-12, maxit=0, thermostat=0, T=298, tau=0.1, nf=1,
                 measure_every=1)
    path = tmp_path / "e.plb"
    with cookio.PlaybackWriter(path, ns) as w:
        w.write(d["L"], d["r"][2])
        for _ in range(3):
            e.md_steps(p, 1)
            w.write(d["L"], e.download_state()[0])
    e.close()
    _, _, _, fr = cookio.read_plb(path)
    ref = np.frombuffer(d["plb"].tobytes()[8:], "<f4").reshape(-1, ns + 1, 3)[1:, 1:]
    ulp = np.spacing(np.abs(ref).astype(np.float32))
    print("max |diff| / ulp %.2f" % float((np.abs(fr - ref) / ulp).max()), flush=True)
    assert np.all(np.abs(fr - ref) <= ulp)
