"""Write symbols.txt: every exported function name of the reference ABI header
with the class of its return type, so the drop-in library can export the same
name set (the reference binding binds every name eagerly at load).

Reads /root/reference/include/atkdsp.h (the ABI header, atkdsp.h:1-41 states
that the header is the ABI). Run once in a container that has the reference;
the output is committed and the build only reads symbols.txt.
"""
import re
import sys
from pathlib import Path

HDR = Path("/root/reference/include/atkdsp.h")
OUT = Path(__file__).with_name("symbols.txt")


def kind(ret: str) -> str:
    r = " ".join(ret.split())
    if r.endswith("*"):
        return "ptr"
    if r == "void":
        return "void"
    if r in ("double", "float"):
        return r
    if r in ("int", "ptrdiff_t", "int64_t"):
        return "neg"          # signed count / status: a stub returns ATKDSP_E_ARG
    if r in ("size_t", "unsigned", "uint32_t", "uint64_t"):
        return "zero"
    raise SystemExit(f"unclassified return type {r!r}")


def main() -> None:
    src = HDR.read_text()
    decls = re.findall(r"ATKDSP_API\s+([^;(]*?)\b(atkdsp_\w+)\s*\(", src)
    seen = {}
    for ret, name in decls:
        seen[name] = kind(ret)
    lines = [f"{n} {k}" for n, k in sorted(seen.items())]
    OUT.write_text("\n".join(lines) + "\n")
    print(f"{len(lines)} symbols -> {OUT}", file=sys.stderr)


if __name__ == "__main__":
    main()
