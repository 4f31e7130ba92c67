"""Generate the golden fixtures of tests/golden/ from the reference build.

Runs oracle/_ref/ref_dump (the unchanged ADPRES Fortran sources compiled by
oracle/Makefile, plus the repo's dump driver) on the reference's sample decks
and on the seeded synthetic cores, and stores inputs and full-precision
outputs as compressed .npz files.  Needs /root/reference and the built
oracle/_ref; the fixtures it writes are what the tests use everywhere else.

    python tests/golden/make_golden.py [name ...]
"""

from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from adpres_amd.configs import generate  # noqa: E402
from adpres_amd.problem import (parse_trace, problem_to_dict, read_ref_dump, write_deck,  # noqa: E402
                                write_state_file, write_xs_override)

REF = "/root/reference/smpl/static"
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")

DECKS = {"iaea3d": "IAEA3Ds", "iaea2d": "IAEA2D", "pnm": "PNM", "biblis": "BIBLIS", "koeberg": "KOEBERG",
         "dvp": "DVP", "adjoint": "adjoint", "fdm": "FDM", "fixed_source": "fixed_source"}


def run(workdir, deck, action, xs="-", state=None):
    args = [DUMP, deck, action, "out.bin", xs] + ([state] if state else [])
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + " ".join(args)], cwd=workdir,
                       capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(os.path.join(workdir, "out.bin")):
        raise RuntimeError(f"ref_dump failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return read_ref_dump(os.path.join(workdir, "out.bin"), action) + (r.stdout,)


def sha(p):
    h = hashlib.sha256()
    for k in ("ix", "iy", "iz", "mat", "xdel", "ydel", "zdel", "sigtr", "siga", "nuf", "sigf", "sigs", "chi", "dc"):
        h.update(np.ascontiguousarray(getattr(p, k)).tobytes())
    return h.hexdigest()


def trace_arrays(stdout, fixed_source=False):
    t = np.asarray(parse_trace(stdout, fixed_source), dtype=np.float64).reshape(-1, 4)
    return dict(trace_p=t[:, 0].astype(np.int32), trace_k=t[:, 1], trace_ser=t[:, 2], trace_fer=t[:, 3])


def save(name, **arrs):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrs)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def deck_solve(name, deck):
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(REF, deck), os.path.join(d, "deck"))
        p, res, out = run(d, "deck", "solve")
    p.name = name
    if p.mode != "FIXEDSRC":
        p.mode = "ADJOINT" if name == "adjoint" else "FORWARD"
    save(name + "_solve", **problem_to_dict(p), Ke=res["Ke"], f0=res["f0"], fs0=res["fs0"], pow=res["pow"],
         dn=res["dn"], df=res["df"], **trace_arrays(out, p.mode == "FIXEDSRC"))
    return p, res


def probes(name, deck, p, res, seed):
    """Nodal-update and BiCGSTAB probes around the converged state."""
    rng = np.random.default_rng(seed)
    N, G = p.nnod, p.ng
    Ke = res["Ke"] * 1.0007
    f0 = res["f0"] * rng.uniform(0.97, 1.03, size=(G, N))
    dn = res["dn"] * rng.uniform(0.9, 1.1, size=(6, G, N)) + rng.uniform(-1e-3, 1e-3, size=(6, G, N))
    b = rng.uniform(0.5, 1.5, size=(G, N))
    nin = 5
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(REF, deck) if deck else deck_path_of(p), os.path.join(d, "deck"))
        write_state_file(os.path.join(d, "st.bin"), nin, Ke, f0, dn, b)
        _, nres, _ = run(d, "deck", "nodal", "-", "st.bin")
        os.remove(os.path.join(d, "out.bin"))
        _, bres, _ = run(d, "deck", "bicg", "-", "st.bin")
    save(name + "_probe", **problem_to_dict(p), Ke=Ke, f0=f0, dn_in=dn, b=b, nin=nin, dn_out=nres["dn"],
         df=nres["df"], ndmax=nres["ndmax"], loc=np.asarray(nres["loc"], np.int32), x_out=bres["x"])


def deck_path_of(p):
    raise RuntimeError("probe without a deck")


def synthetic(name, seed, full=True, workdir=None):
    """workdir: reuse a finished reference run there (deck, out.bin, run.log);
    the 68x68x96 8-group core takes about an hour on one core."""
    gc = generate(name, seed)
    p = gc.problem
    with tempfile.TemporaryDirectory() as d:
        it = (p.nout, p.nin, p.serc, p.ferc, p.nac, p.nupd)
        write_deck(os.path.join(d, "deck"), kern=p.kern, iter_card=it, **gc.deck)
        xs = "-"
        if gc.perturbed:
            write_xs_override(os.path.join(d, "xs.bin"), p)
            xs = "xs.bin"
        if workdir is not None:
            assert open(os.path.join(workdir, "deck")).read() == open(os.path.join(d, "deck")).read()
            q, res = read_ref_dump(os.path.join(workdir, "out.bin"), "solve")
            out = open(os.path.join(workdir, "run.log")).read()
        else:
            q, res, out = run(d, "deck", "solve", xs)
    # the reference must hold exactly the generated problem
    for k in ("ix", "iy", "iz", "mat", "xdel", "ydel", "zdel", "sigtr", "siga", "nuf", "sigf", "sigs", "chi",
              "D", "sigr"):
        assert np.array_equal(getattr(p, k), getattr(q, k)), k
    extra = problem_to_dict(p) if full else {}
    big = p.nnod * p.ng > 1_000_000
    save(f"{name}_s{seed}_solve", **extra, seed=seed, problem_sha=sha(p), Ke=res["Ke"],
         f0=res["f0"][:, ::97] if big else res["f0"], f0_stride=97 if big else 1,
         pow=res["pow"].astype(np.float32) if big else res["pow"],
         dn=res["dn"] if full else res["dn"][:, :, ::97], **trace_arrays(out))


def _batch_one(args):
    """Reference solve of one instance of the batching workload (worker);
    results are cached under `cache` so an interrupted run resumes."""
    seed, i, cache = args
    path = os.path.join(cache, f"{i}.npz")
    if os.path.exists(path):
        z = np.load(path)
        return i, float(z["Ke"]), int(z["niter"]), str(z["sha"]), z["pow"], str(z["msg"])
    from adpres_amd.configs import batch_core
    gc = batch_core(seed, i)
    p = gc.problem
    with tempfile.TemporaryDirectory() as d:
        it = (p.nout, p.nin, p.serc, p.ferc, p.nac, p.nupd)
        write_deck(os.path.join(d, "deck"), kern=p.kern, iter_card=it, **gc.deck)
        write_xs_override(os.path.join(d, "xs.bin"), p)
        try:
            q, res, out = run(d, "deck", "solve", "xs.bin")
        except (RuntimeError, ValueError) as e:
            log = open(os.path.join(d, "deck.out")).read()[-600:] if os.path.exists(os.path.join(d, "deck.out")) else ""
            msg = (str(e)[-600:] + log).replace("\n", " | ")
            np.savez(path, Ke=np.nan, niter=-1, sha=sha(p), pow=np.zeros(0, np.float32), msg=msg)
            return i, float("nan"), -1, sha(p), np.zeros(0, np.float32), msg
    for k in ("ix", "iy", "iz", "mat", "sigtr", "siga", "nuf", "sigf", "sigs", "D", "sigr"):
        assert np.array_equal(getattr(p, k), getattr(q, k)), (i, k)
    n = len(parse_trace(out))
    pw = res["pow"].astype(np.float32)
    np.savez(path, Ke=res["Ke"], niter=n, sha=sha(p), pow=pw, msg="")
    return i, float(res["Ke"]), n, sha(p), pw, ""


def batch(seed=3, n=256, workers=7, sample=8, cache="/tmp/adg_batch_golden"):
    """The batching workload: reference k and outer count of every instance
    (NaN and -1 where the reference stops without a result, with its message
    printed), node powers (float32) of `sample` seeded converged instances."""
    from multiprocessing import Pool
    os.makedirs(cache, exist_ok=True)
    with Pool(workers) as pool:
        out = sorted(pool.map(_batch_one, [(seed, i, cache) for i in range(n)], chunksize=1))
    for o in out:
        if o[2] < 0:
            print(f"core {o[0]}: reference stopped: {o[5][-400:]}")
    ok = np.array([o[2] >= 0 for o in out])
    pick = np.sort(np.random.default_rng(seed).choice(np.nonzero(ok)[0], sample, replace=False))
    arrs = dict(seed=seed, n=n, Ke=np.array([o[1] for o in out]), niter=np.array([o[2] for o in out], np.int32),
                problem_sha=np.array([o[3] for o in out]), pow_cores=pick.astype(np.int32))
    for i in pick:
        arrs[f"pow_{int(i)}"] = out[i][4]
    save(f"batch_s{seed}_solve", **arrs)


SEARCH_LINE = r"^\s*(\d+)\s+(\d+\.\d\d)\s+(\d+\.\d{5})\s+(\S+E[+-]\d\d)\s+(\S+E[+-]\d\d)\s*$"


def bcsearch():
    """The reference program on its critical-boron search deck without T-H
    (smpl/static/CBCsearch: cbsearch, src/mod_th.f90:752-837, calling outer(0)
    after every cross-section update): the printed search table."""
    import re
    exe = os.path.join(ROOT, "oracle", "_ref", "adpres")
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(REF, "CBCsearch"), os.path.join(d, "CBCsearch"))
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} CBCsearch"], cwd=d,
                           capture_output=True, text=True)
    if r.returncode != 0 or "ADPRES EXIT NORMALLY" not in r.stdout:
        raise RuntimeError(r.stdout[-2000:])
    rows = [m.groups() for m in map(re.compile(SEARCH_LINE).match, r.stdout.splitlines()) if m]
    save("cbcsearch_search", n=np.array([int(q[0]) for q in rows], np.int32),
         bcon=np.array([float(q[1]) for q in rows]), k=np.array([float(q[2]) for q in rows]),
         ser=np.array([float(q[3]) for q in rows]), fer=np.array([float(q[4]) for q in rows]))


NEACRP_DECKS = ("A1", "B1", "C1")
NEACRP_HFP = ("A2", "B2", "C2")     # the same searches at hot full power (output sweep only)
NEACRP_FILES = ("neacrp_xsec", "neacrp_geom", "neacrp_cbcs", "neacrp_ftem", "neacrp_mtem", "neacrp_cden")
SEARCH_TH_LINE = (r"^\s*(\d+)\s+(\d+\.\d\d)\s+(\d+\.\d{5})\s+(\S+E[+-]\d\d)\s+(\S+E[+-]\d\d)"
                  r"\s+(\S+E[+-]\d\d)\s*$")


def neacrp_decks(decks):
    """Copies of the NEACRP decks under decks/neacrp/ whose FILE cards name
    the data files beside them (the shipped decks name the author's paths)."""
    out = os.path.join(HERE, "decks", "neacrp")
    os.makedirs(out, exist_ok=True)
    for fn in NEACRP_FILES:
        shutil.copy(os.path.join(REF, "NEACRP", fn), os.path.join(out, fn))
    for deck in decks:
        text = open(os.path.join(REF, "NEACRP", deck)).read().replace("/home/imronuke/ADPRES/smpl/static/NEACRP/",
                                                                         "./")
        with open(os.path.join(out, deck), "w") as fh:
            fh.write(text)


def neacrp():
    """The reference program on its NEACRP critical-boron searches with T-H
    feedback (smpl/static/NEACRP, cases A1, B1, C1: cbsearcht,
    src/mod_th.f90:840-959, whose th_iter calls outer_th between T-H
    updates).  The decks' FILE cards name the author's machine; the copies
    under decks/neacrp/ name the data files beside them.  Stores the printed
    search tables."""
    import re
    exe = os.path.join(ROOT, "oracle", "_ref", "adpres")
    out = os.path.join(HERE, "decks", "neacrp")
    os.makedirs(out, exist_ok=True)
    for fn in NEACRP_FILES:
        shutil.copy(os.path.join(REF, "NEACRP", fn), os.path.join(out, fn))
    neacrp_decks(NEACRP_HFP)
    arrs = {}
    for deck in NEACRP_DECKS:
        text = open(os.path.join(REF, "NEACRP", deck)).read().replace("/home/imronuke/ADPRES/smpl/static/NEACRP/",
                                                                         "./")
        with open(os.path.join(out, deck), "w") as fh:
            fh.write(text)
        with tempfile.TemporaryDirectory() as d:
            for fn in NEACRP_FILES + (deck,):
                shutil.copy(os.path.join(out, fn), os.path.join(d, fn))
            r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} {deck}"], cwd=d,
                               capture_output=True, text=True)
        if r.returncode != 0 or "ADPRES EXIT NORMALLY" not in r.stdout:
            raise RuntimeError(r.stdout[-2000:])
        rows = [m.groups() for m in map(re.compile(SEARCH_TH_LINE).match, r.stdout.splitlines()) if m]
        arrs[deck + "_n"] = np.array([int(q[0]) for q in rows], np.int32)
        arrs[deck + "_bcon"] = np.array([float(q[1]) for q in rows])
        arrs[deck + "_k"] = np.array([float(q[2]) for q in rows])
        arrs[deck + "_therr"] = np.array([float(q[5]) for q in rows])
    save("neacrp_search", **arrs)


MOX_PART1BD = tuple(f"MOX/Part1b/aro{a}" for a in ("A1", "A3", "A5", "A7", "B6", "C3", "C7", "D6", "E5", "E7")) + \
    tuple(f"MOX/Part1d/ari{a}" for a in ("A1", "A3", "A5", "A7", "B6", "C3", "C7", "D6", "E5", "E7"))
SWEEP_DECKS = ("IAEA3Ds", "IAEA2D", "PNM", "BIBLIS", "KOEBERG", "DVP", "FDM", "FDM-1D", "adjoint", "fixed_source",
               "CBCsearch", "MOX/part1_aro_serpent", "MOX/part1_ari_serpent", "MOX/part1_aro_helios",
               "MOX/part1_ari_helios") + MOX_PART1BD


MOX_XTAB_DECKS = ("part2_serpent", "part2_helios", "part3_serpent", "part3_helios")


def mox_xtab():
    """The MOX part-2/3 boron searches with T-H feedback and branch-table
    cross sections (%XTAB, XStab_updt / brInterp, src/mod_xsec.f90:48-84,
    520-788; part 3 with rodded branch tables, crod_tab_updt): the decks
    name their table files at the author's paths; the copies under
    decks/mox/ name the files beside them."""
    import re
    out = os.path.join(HERE, "decks", "mox")
    os.makedirs(out, exist_ok=True)
    xsec = os.path.join(os.path.dirname(REF), "xsec")
    for sub_ in ("SERPENT_CMM", "HELIOS"):
        for fn in os.listdir(os.path.join(xsec, sub_)):
            shutil.copy(os.path.join(xsec, sub_, fn), os.path.join(out, fn))
    for deck in MOX_XTAB_DECKS:
        text = open(os.path.join(REF, "MOX", deck)).read()
        text = re.sub(r"\S*/smpl/xsec/(SERPENT_CMM|HELIOS)/", "./", text)
        with open(os.path.join(out, deck), "w") as fh:
            fh.write(text)


def outputs():
    """The result blocks (power, axial, flux, T-H tables) of the reference
    program's .out file for every self-contained static sample deck, the
    NEACRP searches and the MOX branch-table searches: the regression sweep
    of the drop-in's outputs."""
    sys.path.insert(0, os.path.dirname(HERE))
    from conftest import output_tokens
    exe = os.path.join(ROOT, "oracle", "_ref", "adpres")
    toks = {}
    jobs = [(os.path.basename(d), [os.path.join(REF, d)]) for d in SWEEP_DECKS]
    nea = os.path.join(HERE, "decks", "neacrp")
    neacrp_decks(NEACRP_HFP)
    nea_files = [os.path.join(nea, f) for f in os.listdir(nea)]
    jobs += [(d, nea_files) for d in NEACRP_DECKS + NEACRP_HFP]
    mox_xtab()
    mox = os.path.join(HERE, "decks", "mox")
    jobs += [(d, [os.path.join(mox, f) for f in os.listdir(mox)]) for d in MOX_XTAB_DECKS]
    for deck, files in jobs:
        if deck not in NEACRP_DECKS + NEACRP_HFP + MOX_XTAB_DECKS and not os.path.exists(os.path.join(HERE, "decks",
                                                                                                     deck)):
            shutil.copy(files[0], os.path.join(HERE, "decks", deck))
        with tempfile.TemporaryDirectory() as d:
            for f in files:
                shutil.copy(f, os.path.join(d, os.path.basename(f)))
            r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} {deck}"], cwd=d,
                               capture_output=True, text=True)
            if r.returncode != 0 or "ADPRES EXIT NORMALLY" not in r.stdout:
                raise RuntimeError(deck + r.stdout[-2000:])
            toks[deck] = np.array(output_tokens(open(os.path.join(d, deck + ".out")).read()))
        print(deck, len(toks[deck]))
    save("outputs_sweep", **{k.replace("-", "_"): v for k, v in toks.items()})


TRANS_LINE = r"^\s*(\d+)\s+(\d+\.\d{3})\s+(-?\d+\.\d{4})\s+(\S+E[+-]\d\d)((?:\s+\d+\.\d\d)*)\s*$"


LMW_TIGHT_ITER = "3000 2 1.e-8 1.e-8 13 60 20 60"


def transients():
    """The reference program on its LMW rod-movement transient
    (smpl/transient/LMW: rod_eject, src/mod_trans.f90:15-170, one outer_tr
    per time step): the printed table of time, reactivity, relative power
    and bank positions.  Also for LMW_tight, the same deck with its %ITER
    tolerances tightened from 1e-5 to 1e-8: the deck's own tolerance leaves
    iteration errors that the Crank-Nicolson steps carry as an undamped
    step-to-step oscillation (the reference's tight and loose runs differ by
    2e-3 in relative power), so a comparison of two solvers needs converged
    steps."""
    import re
    exe = os.path.join(ROOT, "oracle", "_ref", "adpres")
    src = os.path.join(os.path.dirname(REF), "transient", "LMW")
    text = open(src).read()
    assert "1200 2 1.e-5 1.e-5 13 60 20 60" in text
    nea = os.path.join(os.path.dirname(REF), "transient", "NEACRP")
    a1t, b1t, c1t = (open(os.path.join(nea, n)).read() for n in ("A1t", "B1t", "C1t"))
    assert "1200 3 1.e-5 1.e-5 15 20 20 80" in a1t and "1200 3 1.e-5 1.e-5 15 20 20 80" in b1t
    assert "1200 3 1.e-5 1.e-5 20 40 20 80" in c1t
    decks = {"LMW": text, "LMW_tight": text.replace("1200 2 1.e-5 1.e-5 13 60 20 60", LMW_TIGHT_ITER),
             # NEACRP rod ejection at hot zero power with T-H feedback (rod_eject_th)
             "A1t": a1t, "A1t_tight": a1t.replace("1200 3 1.e-5 1.e-5 15 20 20 80", "3000 3 1.e-8 1.e-8 15 20 20 80"),
             "B1t_tight": b1t.replace("1200 3 1.e-5 1.e-5 15 20 20 80", "3000 3 1.e-8 1.e-8 15 20 20 80"),
             "C1t_tight": c1t.replace("1200 3 1.e-5 1.e-5 20 40 20 80", "3000 3 1.e-8 1.e-8 20 40 20 80")}
    for name, body in decks.items():
        with open(os.path.join(HERE, "decks", name), "w") as fh:
            fh.write(body)
        with tempfile.TemporaryDirectory() as d:
            shutil.copy(os.path.join(HERE, "decks", name), os.path.join(d, name))
            r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} {name}"], cwd=d,
                               capture_output=True, text=True)
        if r.returncode != 0 or "ADPRES EXIT NORMALLY" not in r.stdout:
            raise RuntimeError(r.stdout[-2000:])
        rows = [m.groups() for m in map(re.compile(TRANS_LINE).match, r.stdout.splitlines()) if m]
        save(name.lower() + "_transient", step=np.array([int(q[0]) for q in rows], np.int32),
             time=np.array([float(q[1]) for q in rows]), rho=np.array([float(q[2]) for q in rows]),
             power=np.array([float(q[3]) for q in rows]),
             banks=np.array([[float(x) for x in q[4].split()] for q in rows]))


TIGHT_ITER = "3000 3 1.e-8 1.e-8 15 20 20 80"


def tight_deck(text):
    """A deck with its %ITER tolerances at 1e-8 (and at most 3000 outers);
    a deck without the card gets the NEACRP decks' card with 1e-8."""
    if "%ITER" not in text:
        return text.rstrip("\n") + "\n\n! Iteration control, tolerances tightened to 1e-8\n%ITER\n" + TIGHT_ITER + "\n"
    lines = text.split("\n")
    i = next(k for k, ln in enumerate(lines) if ln.strip().upper().startswith("%ITER"))
    j = next(k for k in range(i + 1, len(lines)) if lines[k].split("!")[0].strip())
    f = lines[j].split("!")[0].split()
    assert f[2] == "1.e-5" and f[3] == "1.e-5", f
    f[0], f[2], f[3] = "3000", "1.e-8", "1.e-8"
    lines[j] = " ".join(f)
    return "\n".join(lines)


def transients2(stdout_dir=None):
    """The remaining NEACRP rod ejections (A2t, B2t, C2t: hot full power;
    also B1t, C1t, A2t, B2t and C2t as shipped, at their own tolerances)
    and the MOX part-4 rod ejection with branch tables (HELIOS and SERPENT
    sets), each with its %ITER tolerances at 1e-8 (tight_deck): the printed
    tables of time, reactivity, relative power and bank positions.  Each MOX
    deck takes about 20 minutes of reference CPU time; with stdout_dir, a
    reference run's saved standard output <deck>.stdout there is parsed
    instead of running it again."""
    import re
    exe = os.path.join(ROOT, "oracle", "_ref", "adpres")
    tr = os.path.join(os.path.dirname(REF), "transient")
    mox = os.path.join(HERE, "decks", "mox")
    decks = {d + "_tight": tight_deck(open(os.path.join(tr, "NEACRP", d)).read()) for d in ("A2t", "B2t", "C2t")}
    # and the NEACRP decks as shipped (their own %ITER tolerances)
    decks.update({d: open(os.path.join(tr, "NEACRP", d)).read() for d in ("B1t", "C1t", "A2t", "B2t", "C2t")})
    for part in ("part4_helios", "part4_serpent"):
        text = re.sub(r"\S*/smpl/xsec/(SERPENT_CMM|HELIOS)/", "./", open(os.path.join(tr, "MOX", part)).read())
        decks[part + "_tight"] = tight_deck(text)
    for name, body in decks.items():
        where = mox if name.startswith("part4") else os.path.join(HERE, "decks")
        with open(os.path.join(where, name), "w") as fh:
            fh.write(body)
        cached = None
        for sd in (stdout_dir or "").split(os.pathsep):
            if sd and os.path.exists(os.path.join(sd, name + ".stdout")):
                cached = os.path.join(sd, name + ".stdout")
        if cached:
            out = open(cached).read()
        else:
            with tempfile.TemporaryDirectory() as d:
                for fn in ([f for f in os.listdir(mox) if not f.startswith("part")] if where == mox else []):
                    shutil.copy(os.path.join(mox, fn), os.path.join(d, fn))
                shutil.copy(os.path.join(where, name), os.path.join(d, name))
                r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} {name}"], cwd=d,
                                   capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:])
            out = r.stdout
        if "ADPRES EXIT NORMALLY" not in out:
            raise RuntimeError(name + out[-2000:])
        rows = [m.groups() for m in map(re.compile(TRANS_LINE).match, out.splitlines()) if m]
        save(name.lower() + "_transient", step=np.array([int(q[0]) for q in rows], np.int32),
             time=np.array([float(q[1]) for q in rows]), rho=np.array([float(q[2]) for q in rows]),
             power=np.array([float(q[3]) for q in rows]),
             banks=np.array([[float(x) for x in q[4].split()] for q in rows]))


def trstep():
    """One transient step of the reference on the LMW deck (ref_dump trstep:
    the steady state, then rod_eject / trans_calc's set-up with a rod-like
    removal increase and one outer_tr, src/mod_cmfd.f90:800-880): the
    step's inputs (steady state, modified removal, get_exsrc's source and
    dfis) and its results at full precision."""
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(HERE, "decks", "LMW"), os.path.join(d, "deck"))
        p, res, out = run(d, "deck", "trstep")
    p.name = "lmw_trstep"
    p.mode = "RODEJECT"
    keep = {k: np.asarray(v) for k, v in res.items() if k not in ("vdel", "dn", "df")}
    # the step's D-hat and D-tilde as SHA-256 of their float64 bytes (N, G, 6 order)
    for k in ("dn", "df"):
        keep[k + "_sha"] = np.asarray(hashlib.sha256(np.ascontiguousarray(res[k]).tobytes()).hexdigest())
    save("lmw_trstep", **problem_to_dict(p), **keep)


def fwdadj():
    """rod_eject's steady-state pair on LMW_tight (ref_dump fwdadj): outer(0),
    then outer_ad(0) warm-started from it, which keeps the forward D-hats
    (src/mod_cmfd.f90:677-679).  The forward state is stored as hashes (the
    oracle recomputes it bit for bit); the adjoint state in full."""
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(HERE, "decks", "LMW_tight"), os.path.join(d, "deck"))
        p, res, out = run(d, "deck", "fwdadj")
    p.name = "lmw_fwdadj"
    p.mode = "RODEJECT"
    sha = lambda a: np.asarray(hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest())
    save("lmw_fwdadj", **problem_to_dict(p), Ke_s=res["Ke_s"], f0_s_sha=sha(res["f0_s"]),
         fs0_s_sha=sha(res["fs0_s"]), Ke=res["Ke"], f0=res["f0"], fs0=res["fs0"], dn_sha=sha(res["dn"]),
         ndmax=res["ndmax"])


def xsupd():
    """xsec::XS_updt on the NEACRP A1 deck (every feedback table: boron,
    fuel and moderator temperature, coolant density, control rods;
    ref_dump xsupd): the tables, and the node cross sections at the deck's
    parameters and at perturbed node-wise parameters with every bank partly
    inserted.  The first state's outputs are the problem's own arrays."""
    src = os.path.join(HERE, "decks", "neacrp")
    with tempfile.TemporaryDirectory() as d:
        for fn in NEACRP_FILES:
            shutil.copy(os.path.join(src, fn), os.path.join(d, fn))
        shutil.copy(os.path.join(src, "A1"), os.path.join(d, "deck"))
        p, res, out = run(d, "deck", "xsupd")
    p.name = "neacrp_a1_xsupd"
    t = res["tables"]
    arrs = {"t_nmat": t["nmat"], "t_nb": t["nb"], "t_fbmap": t["fbmap"]}
    for k in ("rbcon", "rftem", "rmtem", "rcden", "coreh", "pos0", "ssize"):
        arrs["t_" + k] = t[k]
    for kind in ("base", "boron", "ftem", "mtem", "cden", "rod"):
        for j, x in enumerate(("sigtr", "siga", "nuf", "sigf", "sigs")):
            arrs[f"t_{kind}_{x}"] = t[kind][j]
    s0, s1 = res["states"]
    for i, st in enumerate((s0, s1)):
        for k in ("bcon", "ftem", "mtem", "cden", "bpos"):
            arrs[f"s{i}_{k}"] = st[k]
    for k in ("sigtr", "siga", "nuf", "sigf", "sigs", "D", "sigr"):
        assert np.array_equal(s0[k], getattr(p, k)), k
        arrs["s1_" + k] = s1[k]
    save("neacrp_a1_xsupd", **problem_to_dict(p), **arrs)


def xtab():
    """xsec::XStab_updt on the MOX part-3 HELIOS deck (branch tables with
    one to three points per parameter, materials with and without rodded
    sets, eight rod banks; ref_dump xtab): the flat branch tables, and the
    node cross sections and discontinuity factors at the deck's parameters
    and at perturbed node-wise parameters with every bank part-way in.  The
    first state's outputs are the problem's own arrays."""
    src = os.path.join(HERE, "decks", "mox")
    with tempfile.TemporaryDirectory() as d:
        for fn in os.listdir(src):
            shutil.copy(os.path.join(src, fn), os.path.join(d, fn))
        shutil.copy(os.path.join(src, "part3_helios"), os.path.join(d, "deck"))
        p, res, out = run(d, "deck", "xtab")
    p.name = "mox_part3_helios_xtab"
    b = res["branch"]
    arrs = {"b_" + k: b[k] for k in ("nmat", "nb", "dims", "points", "xsec", "rxsec", "coreh", "pos0", "ssize",
                                     "fbmap")}
    s0, s1 = res["states"]
    for i, st in enumerate((s0, s1)):
        for k in ("bcon", "ftem", "mtem", "cden", "bpos"):
            arrs[f"s{i}_{k}"] = st[k]
    for k in ("sigtr", "siga", "nuf", "sigf", "sigs", "D", "sigr", "dc"):
        assert np.array_equal(s0[k], getattr(p, k)), k
        arrs["s1_" + k] = s1[k]
    save("mox_part3_helios_xtab", **problem_to_dict(p), **arrs)


def adjoint_fdm(seed=1):
    """The 34x34x26-node 2-group core with the finite-difference kernel,
    solved by the reference in %MODE FORWARD and %MODE ADJOINT (outer_ad)
    to 1e-10: both k and the reference's adjoint flux."""
    gc = generate("core3d_2g", seed)
    p = gc.problem
    out = {}
    for mode in ("FORWARD", "ADJOINT"):
        with tempfile.TemporaryDirectory() as d:
            it = (20000, p.nin, 1e-10, 1e-10, 1000, p.nupd)
            write_deck(os.path.join(d, "deck"), kern="FDM", mode=mode, iter_card=it, **gc.deck)
            write_xs_override(os.path.join(d, "xs.bin"), p)
            q, res, log = run(d, "deck", "solve", "xs.bin")
        out[mode] = res
    save(f"core3d_2g_s{seed}_fdm_adjoint", seed=seed, problem_sha=sha(p), Ke=out["FORWARD"]["Ke"],
         Ke_adj=out["ADJOINT"]["Ke"], psi=out["ADJOINT"]["f0"])


def main(names):
    if not os.path.exists(DUMP):
        raise SystemExit("build the reference first: make -C oracle ref")
    for name, deck in DECKS.items():
        if names and name not in names:
            continue
        p, res = deck_solve(name, deck)
        if name in ("iaea3d", "koeberg", "dvp", "pnm"):
            probes(name, deck, p, res, seed=len(name))
    if not names or "core2d_small" in names:
        synthetic("core2d_small", 1)
    if not names or "core3d_2g" in names:
        synthetic("core3d_2g", 1, full=False)
    if "adjoint_fdm" in names:
        adjoint_fdm()
    if not names or "bcsearch" in names:
        bcsearch()
    if not names or "neacrp" in names:
        neacrp()
    if not names or "outputs" in names:
        outputs()
    if not names or "transients" in names:
        transients()
        trstep()
        fwdadj()
    if not names or "xsupd" in names:
        xsupd()
    if not names or "xtab" in names:
        xtab()
    if not names or "transients2" in names:
        transients2(os.environ.get("ADPRES_REF_STDOUT"))
    if "batch" in names:          # 256 reference solves on 7 host cores
        batch()
    if "core3d_8g" in names:      # about an hour of reference CPU time
        synthetic("core3d_8g", 2, full=False, workdir=os.environ.get("CORE3D_8G_RUN"))


if __name__ == "__main__":
    main(sys.argv[1:])
