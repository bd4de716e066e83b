"""Command-line driver: a run from a UDF input, as the reference's main()
(sp_3d_ns.cxx:392-640) for constitutive_eq Navier_Stokes / Electrolyte
with spherical particles placed by INIT_distribution user_specify,
uniform_random, BCC, FCC or random_walk, and Shear_Navier_Stokes_Lees_Edwards
(Time_evolution_hydro_OBL) with or without particles.

    python -m kapsel_old_amd.run -I input.udf -O output.udf -R restart.udf [-D define.udf]

Flags as file_get (input.cxx:1819-1877): -I input, -O output (snapshot
records every GTS steps, Output_udf), -R restart (written after the last
step, Save_Restart_udf), -D accepted and unused (the input's \\include
names its schema).  The time loop is the reference's: steps ts = 0..MSTEP
(MSTEP = GTS * Num_snap), a record before the step whenever ts % GTS == 0,
the restart after step MSTEP.  An input whose resume.Calculation is
CONTINUE (a restart file) resumes: one step from the fresh initial state,
then the saved state replaces it (sp_3d_ns.cxx:546-551, 619-628), and the
loop goes on from the saved ts.

Extra flags (no counterpart in the reference): --state FILE also writes an
exact binary checkpoint (spm_save_state) after the last step; --resume-state
FILE starts from one instead of the input's resume data; --fp32 runs the
FP32 path; --schedule literal selects the per-routine schedule; --stress
FILE (Lees-Edwards input) writes Mean_shear_stress's line after every step,
as the reference prints it to stderr (sp_3d_ns.h:209-358).

The context counts ts after each step (spm_step), the reference's jikan.ts
is the index of the step running; so the restart holds ts = (steps done) - 1
and a restored context is given ts + 1.
"""
import argparse
import sys

import numpy as np

from . import abi, udf
from .spm import SPM, SCHEDULE_FUSED, SCHEDULE_LITERAL


def parse_args(argv):
    # the reference's flags take the value attached (-Iinput.udf); a
    # separate argument is accepted as well
    norm = []
    for a in argv:
        if len(a) > 2 and a[0] == "-" and a[1] in "IODRSM" and a[2] != "-":
            norm += [a[:2], a[2:]]
        else:
            norm.append(a)
    ap = argparse.ArgumentParser(prog="python -m kapsel_old_amd.run")
    ap.add_argument("-I", dest="input", required=True)
    ap.add_argument("-O", dest="output", required=True)
    ap.add_argument("-R", dest="restart", required=True)
    ap.add_argument("-D", dest="define")
    ap.add_argument("-S", dest="summary")
    ap.add_argument("-M", dest="control")
    ap.add_argument("--state")
    ap.add_argument("--resume-state")
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--schedule", choices=["fused", "literal"], default="fused")
    ap.add_argument("--stress")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args(norm)


def _g(x):
    """printf's %g"""
    return "%g" % x


def initial_particles(u, prm, run, resumed):
    """Init_Particle and the MT19937 state it leaves (None: the context's
    own seed-0 state); for a resumed run of another distribution only the
    species matter (the restart replaces the rest)."""
    if udf.DISTRIBUTIONS[run.distribution] in ("user_specify", "uniform_random", "BCC", "FCC", "random_walk"):
        p, mt, mti = u.init_particles(rng=True)
        return p, (mt, mti)
    if not resumed:
        raise SystemExit("INIT_distribution %s is not run by this driver" % udf.DISTRIBUTIONS[run.distribution])
    p = abi.new_particles(prm.n_particles)
    counts = [u.get_int("object_type.spherical_particle.Particle_spec[%d].Particle_number" % s)
              for s in range(prm.n_species)]
    p["spec"] = np.repeat(np.arange(prm.n_species), counts)
    L = np.array([prm.nx, prm.ny, prm.nz]) * prm.dx
    p["x"] = (np.arange(prm.n_particles)[:, None] + 0.5) / max(prm.n_particles, 1) * L  # placeholders
    p["x_nopbc"] = p["x"]
    return p, None


def restore(ctx, u, prm, n_ion, sheared, log):
    """Force_restore_parameters into the context.  Under Lees-Edwards shear
    only degree_oblique is restored (resume.cxx:335-339); Shear_rate_eff
    stays as the first step measured it."""
    r = udf.read_restart(u, (prm.nx, prm.ny, prm.nz), prm.n_particles, n_ion, particles=ctx.get_particles())
    if sheared:
        ctx.put_oblique(r["degree_oblique"], ctx.get_oblique()[1])
    ctx.put_field("zeta0", r["zeta0"])
    ctx.put_field("zeta1", r["zeta1"])
    ctx.put_uk_dc(r["uk_dc"])
    ctx.put_particles(r["particles"])
    for n in range(n_ion):
        ctx.put_concentration(n, r["concentration"][n])
    ctx.put_time(r["ts"] + 1, r["time"])
    log("############################ Parameters are restored.")
    return r["ts"]


def main(argv=None):
    a = parse_args(sys.argv[1:] if argv is None else argv)
    log = (lambda *m: None) if a.quiet else (lambda *m: print(*m, file=sys.stderr, flush=True))
    u = udf.UDF(a.input)
    prm, el, run = u.params()
    if el is not None and run.poisson_boltzmann and not (el.n_spec == 2 and prm.n_particles > 1):
        raise SystemExit("INIT_profile Poisson_Boltzmann is run for salt with several particles only")
    sheared = run.equation == udf.LEES_EDWARDS
    if a.fp32:
        prm.precision = abi.FP32
    gts, mstep = run.gts, run.gts * run.num_snap
    resumed = u.get_string("resume.Calculation") == "CONTINUE"
    write_udf = u.get_string("output.UDF") == "ON"
    n_ion = el.n_spec if el is not None else 0
    log("#mesh = %d * %d * %d, particles %d, DT %.17g, MSTEP %d, GTS %d" %
        (prm.nx, prm.ny, prm.nz, prm.n_particles, prm.dt, mstep, gts))

    ctx = SPM(prm)
    ctx.set_schedule(SCHEDULE_LITERAL if a.schedule == "literal" else SCHEDULE_FUSED)
    p, rng = initial_particles(u, prm, run, resumed or a.resume_state)
    ctx.put_particles(p)
    if rng is not None:  # the thermostat goes on from the state the initialisation left
        ctx.put_rng(*rng)
    ctx.init_fluid()
    if el is not None:
        ctx.set_electrolyte(el)
        if run.poisson_boltzmann:
            log("# Poisson-Boltzmann start: %d iterations" % ctx.Init_rho_ion_PB())
        else:
            ctx.Init_rho_ion()
    if sheared:
        ctx.set_lees_edwards(run.shear_rate)
    if write_udf:
        udf.create_output(a.input, a.output)

    ts = 0
    if a.resume_state:
        ctx.load_state(a.resume_state)
        ts = ctx.get_time()[0]
    elif resumed:
        ctx.step(1)  # the first loop iteration runs before the restore
        ts = restore(ctx, u, prm, n_ion, sheared, log) + 1
    fstress = None
    if a.stress:
        if not sheared:
            raise SystemExit("--stress needs Lees-Edwards input")
        fstress = open(a.stress, "w")
        fstress.write("#1:time 2:shear_rate 3:shear_strain_temporal 4:lj_dev_stress_temporal "
                      "5:shear_stress_temporal_old 6:shear_stress_temporal_new 7:viscosity \n")
    while ts <= mstep:
        if ts % gts == 0:
            if write_udf:
                udf.append_record(a.output, ts, ctx.get_particles())
            log("# Step: %9d/%9d" % (ts, mstep))
        n = 1 if fstress else min((ts // gts + 1) * gts, mstep + 1) - ts
        ctx.step(n)
        ts += n
        if fstress:
            st = ctx.le_shear_stress()
            fstress.write(" ".join(_g(st[k]) for k in abi.LE_STRESS_KEYS) + "\n")
    if fstress:
        fstress.close()
    ts_done, time = ctx.get_time()
    shape = (prm.nx, prm.ny, prm.nz)
    conc = [ctx.get_concentration(n) for n in range(n_ion)] if n_ion else None
    udf.write_restart(a.input, a.restart, shape, ctx.get_field("zeta0"), ctx.get_field("zeta1"), ctx.get_uk_dc(),
                      ctx.get_particles(), ts_done - 1, time, conc=conc,
                      degree_oblique=ctx.get_oblique()[0] if sheared else 0.0)
    if a.state:
        ctx.save_state(a.state)
    ctx.close()
    u.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
