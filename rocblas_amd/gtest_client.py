"""rocblas-test for the trsm path: runs the reference's gtest case matrix
(clients/gtest/trsm_gtest.yaml, or the same structure as JSON) against this
library, case by case, with the reference's procedures and tolerances
(rocblas_amd.clients).

    python -m rocblas_amd.gtest_client --yaml clients/gtest/trsm_gtest.yaml \\
        -I clients/include --category quick

Cases this library does not run are counted as skipped with the reason:
complex precisions, HipManagedMemory, multi-thread / multi-device cases.
The exit status is non-zero when a case fails.
"""
import argparse
import sys
import time

from . import Handle
from . import clients as C


def select(cases, categories=(), name_filter="", function_filter="", max_bytes=0):
    """(runnable Arguments, {reason: count} of the skipped) for the given filters."""
    run, skipped = [], {}
    for case in C.dedupe(cases):
        if categories and case.get("category") not in categories:
            continue
        if name_filter and name_filter not in str(case.get("name", "")):
            continue
        if function_filter and function_filter not in str(case.get("function", "")):
            continue
        kind = C.case_kind(case)
        if kind == "run":
            arg = C.Arguments.from_case(case)
            if max_bytes and C.case_bytes(arg) > max_bytes:
                kind = f"needs more than {max_bytes} bytes of device memory"
            else:
                run.append(arg)
                continue
        skipped[kind] = skipped.get(kind, 0) + 1
    return run, skipped


def run(args, handle, verbose=False, out=None, progress=0):
    """Run the cases; returns the list of (label, message) failures.  With
    progress > 0 a line is printed every that many cases."""
    out = out or sys.stdout
    failures = []
    for i, arg in enumerate(args):
        if progress and i and i % progress == 0:
            print(f"[ PROGRESS ] {i}/{len(args)}, {len(failures)} failed", file=out, flush=True)
        arg.seed = i
        label = arg.label()
        # a large case can take a while on the host: name it before it starts
        if verbose or C.case_bytes(arg) > (1 << 30) or arg.K > 8192:
            print(f"[ RUN      ] {label}", file=out, flush=True)
        try:
            C.run_case(handle, arg)
        except C.CaseFailure as e:
            failures.append((label, str(e)))
            print(f"[  FAILED  ] {label}: {e}", file=out, flush=True)
            continue
        if verbose:
            print(f"[       OK ] {label}", file=out, flush=True)
    return failures


def main(argv=None):
    p = argparse.ArgumentParser(prog="rocblas-test (trsm)", description=__doc__,
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    src = p.add_mutually_exclusive_group(required=True)
    src.add_argument("--yaml", help="gtest-format YAML file")
    src.add_argument("--json", help="the same structure as JSON")
    p.add_argument("-I", "--include", action="append", default=[], help="YAML include directory")
    p.add_argument("--category", default="quick",
                   help="comma-separated categories (quick, pre_checkin, nightly, stress); "
                        "'all' for every one")
    p.add_argument("--name_filter", default="")
    p.add_argument("--function_filter", default="")
    p.add_argument("--max_bytes", type=int, default=32 << 30,
                   help="skip cases whose operands need more device memory")
    p.add_argument("-v", "--verbose", action="store_true")
    ns = p.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("rocblas-test: no GPU")
    cases = C.cases_from(ns.yaml or ns.json, ns.include)
    cats = () if ns.category == "all" else tuple(ns.category.split(","))
    args, skipped = select(cases, cats, ns.name_filter, ns.function_filter, ns.max_bytes)
    print(f"[==========] Running {len(args)} tests ({ns.category}).", flush=True)
    t0 = time.time()
    with Handle() as h:
        failures = run(args, h.ptr, ns.verbose, progress=500)
    print(f"[==========] {len(args)} tests ran. ({time.time() - t0:.1f} s total)")
    print(f"[  PASSED  ] {len(args) - len(failures)} tests.")
    for reason, n in sorted(skipped.items()):
        print(f"[  SKIPPED ] {n} tests: {reason}")
    if failures:
        print(f"[  FAILED  ] {len(failures)} tests.")
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
