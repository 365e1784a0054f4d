"""Writes rapid_mix.json: input/output vectors of the reference's own rapid_mix
(include/sqsgen/core/helpers/rapidhash.h:149-212, built into oracle/_ref by oracle/Makefile)
and of the rapidrand draw stream it drives (core/shuffle.h:23-26) for a few seeds.

Run from the repository root after `make -C oracle`: python tests/golden/make_rapid_mix.py
"""
import json
import random
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

M64 = (1 << 64) - 1


def main():
    ref = O.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (the reference tree is needed)")
    rnd = random.Random(20261203)
    pairs = [(0, 0), (1, 1), (M64, M64), (M64, 2), (1 << 63, 3)]
    pairs += [(rnd.getrandbits(64), rnd.getrandbits(64)) for _ in range(200)]
    mix = [[a, b, ref.ref_rapid_mix(a, b)] for a, b in pairs]
    s0, s1 = ref.ref_rapid_secret(0), ref.ref_rapid_secret(1)
    streams = {}
    for seed in (0, 42, 7, 1001, 0xDEADBEEFCAFEBABE):
        st, out = seed, []
        for _ in range(64):
            st = (st + s0) & M64
            out.append(ref.ref_rapid_mix(st, st ^ s1))
        streams[str(seed)] = out
    doc = dict(source="include/sqsgen/core/helpers/rapidhash.h (compiled as oracle/_ref)",
               secret0=s0, secret1=s1, mix=mix, rapidrand_streams=streams)
    Path(__file__).with_name("rapid_mix.json").write_text(json.dumps(doc))
    print(f"{len(mix)} mix vectors, {len(streams)} streams")


if __name__ == "__main__":
    main()
