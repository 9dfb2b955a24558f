"""Plain device-to-device copy of a known size (3 x 1 GiB), run under
`rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` to learn what those
counters read per byte really moved; the HBM traffic figures of the cast
kernel are scaled by the result.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <out> -o run -- \\
        python -m hptracing_amd.copy_calibration
"""
import torch


def main() -> None:
    n = 1 << 28  # 256 Mi floats = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 3 x 1 GiB", flush=True)


if __name__ == "__main__":
    main()
