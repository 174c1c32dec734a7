"""The unmeasured MD step replayed from a captured hipGraph gives the same
bits as the eager launches: every call replays the graph several times,
with eager measured steps, state downloads and host synchronizations
between the replays, and each round is compared with an engine that
launches every step eagerly."""
import numpy as np
import pytest

import oracle
from macsimus_amd import abi, workloads
from macsimus_amd.engine import Engine
from macsimus_amd.system import engine_provider

pytestmark = pytest.mark.gpu


def run_pair(bs, r, r1, p, mode, rounds=5, per=7, aspc=None):
    engs = []
    for graph in (True, False):
        e = Engine(bs, mode)
        e.set_graph(graph)
        e.upload_state(r, r1)
        if aspc is not None:
            e.set_aspc(*aspc)
        engs.append(e)
    for k in range(rounds):
        outs = [e.md_steps(p, per) for e in engs]
        sg, se = [e.download_state() for e in engs]
        assert np.array_equal(sg[0], se[0]) and np.array_equal(sg[1], se[1]), "round %d" % k
        assert outs[0].frc.En_el == outs[1].frc.En_el and outs[0].En_kin == outs[1].En_kin
    for e in engs:
        e.close()


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_graph_replay_matches_eager(mode):
    d, meta = oracle.load_fixture("spce1000_md")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    nf = workloads.degrees_of_freedom(meta["forcefield"], meta["counts"])
    p = abi.Step(h=0.002, eps=1e-8, maxit=0, thermostat=1, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=0)
    run_pair(bs, d["r"], np.zeros_like(d["r"]), p, mode)


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_graph_replay_matches_eager_drude(mode):
    """Drudes: the first steps ramp the predictor eagerly, then the graph"""
    d, meta = oracle.load_fixture("polar306_aspc")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    r = np.concatenate([d["r"][1], d["rpol"][1]])
    r1 = np.concatenate([d["r"][1] - d["r"][0], d["rpol"][1]])
    p = abi.Step(h=meta["h"], eps=1e-10, maxit=0, thermostat=0, T=298.0, tau=0.1, nf=1.0,
                 measure_every=0, aspc_k=oracle.ASPC_K, scf_iter=1, scf_omega=0.0)
    run_pair(bs, r, r1, p, mode, rounds=4, per=6, aspc=(oracle.ASPC_K, 0, None))


@pytest.mark.parametrize("mode", [abi.MDX_FP64, abi.MDX_MIXED])
def test_kspace_overlap_gives_the_same_bits(mode):
    """k-space on its own stream beside the pair kernel (default) and all
    on one stream: the same forces, energies and MD trajectory, bit for
    bit, eager and graph-replayed"""
    d, meta = oracle.load_fixture("spce1000_md")
    bs = oracle.fixture_system(meta, provider_=engine_provider())
    nf = workloads.degrees_of_freedom(meta["forcefield"], meta["counts"])
    p = abi.Step(h=0.002, eps=1e-8, maxit=0, thermostat=1, T=298.0, tau=0.1, nf=float(nf),
                 measure_every=0)
    engs = []
    for ov in (True, False):
        e = Engine(bs, mode)
        e.set_overlap(ov)
        engs.append(e)
    (f0, o0), (f1, o1) = [e.forces(d["r"], 1) for e in engs]
    assert np.array_equal(f0, f1) and o0.En_el == o1.En_el and list(o0.Pvir) == list(o1.Pvir)
    for e in engs:
        e.upload_state(d["r"], np.zeros_like(d["r"]))
    for k in range(4):
        outs = [e.md_steps(p, 6) for e in engs]
        s0, s1 = [e.download_state() for e in engs]
        assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[1], s1[1]), k
        assert outs[0].frc.En_el == outs[1].frc.En_el and outs[0].En_kin == outs[1].En_kin
    for e in engs:
        e.close()
