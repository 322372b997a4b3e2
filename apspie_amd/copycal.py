"""Plain device-to-device copy of a known size through the library: the
reference point for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters to
bytes (run under rocprofv3 --pmc; the copy kernel reads and writes exactly
`nbytes` per dispatch).

    python3 -m apspie_amd.copycal [nbytes] [reps]
"""
import sys

import apspie_amd as A


def main(nbytes=1 << 30, reps=3):
    with A.Context(0) as ctx:
        src = ctx.dev_alloc(nbytes)
        dst = ctx.dev_alloc(nbytes)
        for _ in range(reps):
            A._check(A.lib().apspie_memcpy(ctx.h, dst.ptr, src.ptr, nbytes, 2), ctx.h)
        ctx.sync()
        src.free()
        dst.free()
    print(f"copied {nbytes} bytes x{reps} (device to device)", flush=True)


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:3]))
