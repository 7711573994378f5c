"""Per-kernel resource table from the compiler's kernel-resource-usage
remarks (`make resources` pipes them in): VGPRs, spills, scratch, LDS and
the occupancy the compiler computes for each kernel of the library."""
import re
import subprocess
import sys

FIELDS = {"VGPRs": "vgpr", "AGPRs": "agpr", "ScratchSize [bytes/lane]": "scratch",
          "Occupancy [waves/SIMD]": "waves", "LDS Size [bytes/block]": "lds",
          "VGPRs Spill": "vspill", "SGPRs Spill": "sspill", "TotalSGPRs": "sgpr"}


def main():
    rows, cur = {}, None
    for line in sys.stdin:
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            cur = rows.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+([A-Za-z][^:\[]*(?:\[[^\]]*\])?):\s+(\S+)", line)
        if m and cur is not None and m.group(1).strip() in FIELDS:
            cur[FIELDS[m.group(1).strip()]] = m.group(2)
    names = sorted(rows)
    try:
        out = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True,
                             check=True).stdout.splitlines()
        pretty = dict(zip(names, out))
    except (OSError, subprocess.CalledProcessError):
        pretty = {n: n for n in names}
    print("| Kernel | VGPRs | AGPRs | SGPRs | VGPR spill | scratch B/lane | LDS B/block | waves/SIMD |")
    print("|---|---|---|---|---|---|---|---|")
    for n in sorted(names, key=lambda n: pretty[n]):
        r = rows[n]
        k = re.sub(r"^void le2d::", "", pretty[n]).split("(")[0]
        print(f"| `{k}` | {r.get('vgpr', '')} | {r.get('agpr', '')} | {r.get('sgpr', '')} | "
              f"{r.get('vspill', '')} | {r.get('scratch', '')} | {r.get('lds', '')} | "
              f"{r.get('waves', '')} |")


if __name__ == "__main__":
    main()
