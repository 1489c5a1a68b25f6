#!/usr/bin/env python3
"""Writes tests/golden/dumper_threads.json: what the reference dumpers, compiled from their sources
(oracle/Makefile target ref -> oracle/_ref), print for the inputs of
tests/test_dumpers.py::test_threaded_counts_and_masks_like_reference -- exit status, stderr, record
count and SHA-256 of the record set (the threaded I stream sorted by (key, id), since its threads
interleave in whatever order they finish; the D streams as written). Run where oracle/_ref is built."""
import hashlib
import json
import os
import subprocess
import sys
from math import prod

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "carmichael-frontier_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(HERE)]

import oracle as O  # noqa: E402
import test_dumpers as T  # noqa: E402


def run(prog, text, env):
    pr = subprocess.run([os.path.join(O.REF_DIR, prog)], input=text.encode(), capture_output=True,
                        env=dict(os.environ, **env), timeout=600, check=False)
    return pr.returncode, pr.stdout, pr.stderr.decode()


def main():
    out = {}
    text, env = T.threads_case()
    st, data, err = run("mitm64_idump_stream2", text, env)
    a = np.frombuffer(data, np.uint64).reshape(-1, 2)
    a = np.ascontiguousarray(a[np.lexsort((a[:, 1], a[:, 0]))])
    out["idump_threads"] = {"status": st, "stderr": err, "records": int(a.shape[0]),
                            "sorted_sha": hashlib.sha256(a.tobytes()).hexdigest()}
    out["ddump_masks"] = []
    for r, text, env in T.mask_cases():
        st, data, err = run("mitm64_ddump", text, env)
        out["ddump_masks"].append({"r": r, "status": st, "stderr": err, "bytes": len(data),
                                   "sha": hashlib.sha256(data).hexdigest()})
    with open(os.path.join(HERE, "dumper_threads.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
