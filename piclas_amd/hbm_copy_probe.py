"""Plain device-to-device copy of a known size.

Used to calibrate the rocprofv3 HBM counters: run it under
``rocprofv3 --pmc FETCH_SIZE`` (and ``WRITE_SIZE``) and divide the known bytes
(read = written = NBYTES per copy) by what the counter reports per dispatch.
The resulting factors scale the same counters read for the stage's kernels.

  python -m piclas_amd.hbm_copy_probe [NBYTES]     (default 1 GiB)
"""
import sys

import torch


def main():
    nbytes = int(sys.argv[1]) if len(sys.argv) > 1 else (1 << 30)
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d bytes x3 (%d KiB per copy read and written)" % (nbytes, nbytes // 1024), flush=True)


if __name__ == "__main__":
    main()
