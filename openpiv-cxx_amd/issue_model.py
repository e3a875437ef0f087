"""VALU issue model of a kernel: the sum, over the VALU instructions of its
gfx950 assembly, of each opcode's measured issue cost (lib/opiv_issue_probe),
and the kernel time that sum implies when every SIMD issues without a gap.

    hipcc --offload-arch=gfx950 -O3 -std=c++17 -I../include -Icsrc \
          -Xclang -target-feature -Xclang -packed-fp32-ops \
          --cuda-device-only -S csrc/corr_fft.hip -o corr_fft.s
    python3 issue_model.py corr_fft.s probe.txt corr_fft_kernelILi32ELi2ELb0ELi2ELi256ELi0E --waves 16260

`probe.txt` is the output of lib/opiv_issue_probe on the GPU.  The kernels
are straight-line code (every loop over rows, windows and FFT stages is
unrolled at compile time), so the static count is the count each wave
issues, less the branches most waves skip (SQ_INSTS_VALU per wave reads
3767 against a static 4016 at 32 x 32).  MFMA instructions run
on the matrix pipe and are not counted.
"""
import argparse
import collections
import re
import sys

# probe row for each opcode family (the probe measures one representative)
FAMILY = [
    (r"_dpp|dpp", "v_mov_b32_dpp"),
    (r"^v_(add|sub|subrev)_f32_e64$", "v_add_f32_e64 (VOP3)"),
    (r"^v_(add|sub|subrev)_f32", "v_add_f32_e32"),
    (r"^v_mul_f32", "v_mul_f32_e32"),
    (r"^v_fma_f32$", "v_fma_f32 (VOP3)"),
    (r"^v_(fmac|fmamk|fmaak)_f32", "v_fmac_f32_e32"),
    (r"^v_mov_b(32|64)", "v_mov_b32_e32"),
    (r"^v_(add|sub|subrev)_(u32|co_u32|i32)", "v_add_u32_e32"),
    (r"^v_(and|or|not)_b32", "v_and_b32_e32"),
    (r"^v_xor_b32", "v_xor_b32_e32"),
    (r"^v_(lshlrev|lshrrev|ashrrev)_", "v_ashrrev_i32_e32"),
    (r"^v_(max|min)_f32", "v_max_f32_e32"),
    (r"^v_(max3|min3)_", "v_max3_f32"),
    (r"^v_med3_", "v_med3_f32"),
    (r"^v_(max|min)_[iu]32", "v_max_i32_e32"),
    (r"^v_(mul_lo|mul_hi|mad_u32|mul_u32|mad_i32|mul_i32)", "v_mul_lo_u32"),
    (r"^v_cvt_", "v_cvt_f32_i32_e32"),
    (r"^v_(cmp|cmpx|cndmask)", "v_cmp_gt_f32 + v_cndmask_b32"),
    (r"^v_(fma|mul)_f64|^v_fmac_f64", "v_fma_f64"),
    (r"_f64", "v_add_f64"),
    (r"^v_", "v_bfi_b32"),          # other VOP3 integer/bit operations
]


def load_costs(path):
    costs = {}
    for line in open(path):
        m = re.match(r"^(\S.*?)\s{2,}[\d.]+ ms\s+([\d.]+) cycles", line)
        if m:
            costs[m.group(1).strip()] = float(m.group(2))
    if not costs:
        sys.exit(f"no probe rows in {path}")
    return costs


def classify(op):
    for pat, row in FAMILY:
        if re.search(pat, op):
            return row
    return None


def kernel_body(asm_lines, pattern):
    """lines of the first kernel whose label contains `pattern`"""
    for i, line in enumerate(asm_lines):
        label = line.split(":")[0]
        if line.startswith("_Z") and ":" in line and pattern in label:
            j = i
            while j < len(asm_lines) and not asm_lines[j].startswith(".Lfunc_end"):
                j += 1
            return label, asm_lines[i:j]
    sys.exit(f"no kernel matching {pattern}")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("asm")
    ap.add_argument("probe")
    ap.add_argument("kernel", help="a substring of the kernel's mangled name")
    ap.add_argument("--waves", type=int, required=True, help="waves the launch runs")
    ap.add_argument("--simds", type=int, default=1024)
    ap.add_argument("--clock-ghz", type=float, default=2.4, help="the clock the probe's cycles refer to")
    args = ap.parse_args()
    costs = load_costs(args.probe)
    name, body = kernel_body(open(args.asm).read().split("\n"), args.kernel)
    units = collections.Counter()
    count = collections.Counter()
    for l in body:
        t = l.strip().split()
        if not t or not t[0].startswith("v_") or t[0].startswith(("v_mfma", "v_smfmac")):
            continue
        row = classify(t[0] + (" dpp" if "dpp" in l else ""))
        units[row] += costs[row]        # (the cmp + cndmask row is per instruction of the pair)
        count[row] += 1
    total = sum(units.values())
    waves_per_simd = args.waves / args.simds
    t_us = total * waves_per_simd / (args.clock_ghz * 1e3)
    print(name)
    print(f"VALU instructions per wave {sum(count.values())}, issue cycles per wave {total:.0f}, "
          f"{waves_per_simd:.2f} waves per SIMD -> {t_us:.1f} us at full issue")
    for row, u in units.most_common():
        print(f"  {row:32s} n={count[row]:5d}  cycles={u:7.0f}  {100 * u / total:4.1f}%")


if __name__ == "__main__":
    main()
