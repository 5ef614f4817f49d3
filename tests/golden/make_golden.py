"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in a container that has the reference sources (it builds
oracle/_ref/matched_filter_CPU.so with `make -C oracle ref`, i.e. the
reference's matched_filter.c compiled with its own flags).  Inputs are
regenerated by the tests from the seeds recorded here and checked against the
stored SHA-256 digests; the outputs are the reference's own float32 bits.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from oracle import cpu  # noqa: E402
from tests.golden import cases  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "ref"])
    meta = dict(host_cpu=cpu_model(), flags="-O3 -fopenmp -fPIC -ftree-vectorize -march=native -std=gnu99")
    out = {}
    for name, build in cases.CASES.items():
        inp = build()
        for key, opts in cases.VARIANTS[name].items():
            opts = dict(opts)
            if opts.get("kind") == "variable_precise":
                opts["n_samples_template"] = inp["lengths"]
            cc = cpu.run(inp["templates"], inp["moveouts"], inp["weights"],
                         inp["data"], inp["step"], lib="ref", **opts)
            tag = name + "__" + key
            out[tag + "__sha"] = np.array(sha(cc))
            out[tag + "__shape"] = np.array(cc.shape)
            if cc.size <= 200_000:
                out[tag + "__full"] = cc
            else:
                flat = cc.reshape(cc.shape[0], -1)
                out[tag + "__every97"] = flat[:, ::97]
                out[tag + "__max"] = flat.max(axis=1)
                out[tag + "__argmax"] = flat.argmax(axis=1)
                out[tag + "__zeros"] = (flat == 0).sum(axis=1)
                out[tag + "__sum64"] = flat.astype(np.float64).sum(axis=1)
                wins = []
                for t, lag in inp.get("events", []):
                    o = lag // inp["step"]
                    wins.append(flat[t, max(0, o - 64):o + 65])
                if wins:
                    out[tag + "__event_windows"] = np.stack(
                        [np.pad(w, (0, 129 - len(w))) for w in wins])
        out[name + "__input_sha"] = np.array(cases.input_digest(inp))
    out["meta_host_cpu"] = np.array(meta["host_cpu"])
    out["meta_flags"] = np.array(meta["flags"])
    path = os.path.join(HERE, "reference_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
