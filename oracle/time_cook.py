"""Times the reference cook (threaded build, PARALLEL=1 with barriers,
cook's own gcc options) on this machine's cores: the 1000 SPC/E water box
(3,000 sites, L = 31.04 A, 9 A cutoff, 1000 steps) and a short sample of
the 194,304-site brine (64,000 SPC/E + 1,152 NaCl, 12 A cutoff), both with
Ewald at el.epsr = el.epsk = 0.05, SHAKE, leap-frog h = 2 fs and a 298 K
Berendsen thermostat from cook's crystal start.  The throughput is cook's
own `speed = ... ns/day` line (sim/main.c:1978-1984); the threads are
NSLOTS (at most 24, sim/siminit.c:1304).

    python oracle/time_cook.py [--threads N] [--c2-steps K] [--c1-steps K]

Builds oracle/_ref/cookewslcP1 from /root/reference with
oracle/Makefile.ref (target perf) first.  Prints one JSON line."""
import argparse
import json
import os
import platform
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
sys.path.insert(0, REPO)
import make_golden as mg  # noqa: E402

COOKP = os.path.join(REPO, "oracle", "_ref", "cookewslcP1")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_case(species, types, counts, L, cutoff, steps, noint, threads):
    with tempfile.TemporaryDirectory() as td:
        mg.write_ble(os.path.join(td, "sys.ble"), species, types)
        lines = ["N[%d]=%d" % (i, n) for i, n in enumerate(counts)]
        lines += ["L[0]=%.10g L[1]=%.10g L[2]=%.10g" % (L, L, L),
                  "cutoff=%g" % cutoff, "LJcutoff=cutoff", "el.epsr=0.05", "el.epsk=0.05",
                  "h=0.002", "noint=%d" % noint, "T=298", "epsc=1e-7", ";"]
        open(os.path.join(td, "sys.def"), "w").write("\n".join(lines) + "\n")
        open(os.path.join(td, "md.get"), "w").write(
            'init="crystal"\nthermostat="Berendsen"\ntau.T=0.1\nno=%d;\n' % (steps // noint))
        env = dict(os.environ, NSLOTS=str(threads))
        t0 = time.perf_counter()
        p = subprocess.run([COOKP, "sys", "md"], cwd=td, env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=3600)
        wall = time.perf_counter() - t0
        txt = open(os.path.join(td, "md.prt")).read() if os.path.exists(
            os.path.join(td, "md.prt")) else ""
        if p.returncode != 0:
            sys.stderr.write(p.stdout.decode(errors="replace")[-3000:])
            raise RuntimeError("cook failed")
    m = re.findall(r"speed = \S+ cycles/min = \S+ cycles/hour = (\S+) ns/day", txt)
    d = re.findall(r"duration (\S+) s", txt)
    th = re.findall(r"PARALLEL=1 \(linked-cell list parallelized\), (\d+) threads", txt)
    nc = re.findall(r"No\.cell=\[?(\d+)", txt)
    return dict(steps=steps, ns_day=float(m[-1]), duration_s=float(d[-1]), wall_s=wall,
                threads=int(th[-1]) if th else None, ncell0=int(nc[-1]) if nc else None,
                ms_per_step=float(d[-1]) / steps * 1e3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=min(os.cpu_count() or 1, 24))
    ap.add_argument("--c1-steps", type=int, default=1000)
    ap.add_argument("--c2-steps", type=int, default=20)
    a = ap.parse_args()
    subprocess.run(["make", "-s", "-f", "oracle/Makefile.ref", "-j8", "perf"], cwd=REPO, check=True)
    from macsimus_amd import workloads
    spce = mg.spce_species()
    out = dict(cpu=cpu_model(), threads=a.threads, build="PARALLEL=1 BARRIER, gcc -O2 -ffast-math")
    out["water_3000"] = run_case([spce], ["H", "O"], [1000], 31.04, 9.0, a.c1_steps, 10, a.threads)
    print(json.dumps(out), file=sys.stderr, flush=True)
    wl = workloads.brine_c2()
    out["brine_194304"] = run_case([mg.ion_species("NA", 1.0), mg.ion_species("CL", -1.0), spce],
                                   ["H", "O", "NA", "CL"], wl["counts"], wl["L"], wl["cutoff"],
                                   a.c2_steps, 1, a.threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
