"""Writes loader_symbols.json: the symbols the reference's Python loader
resolves when it loads the library, and the ABI version it requires.

Source: python/quantkernel/_loader.py:91-316 (the _FUNCTION_SIGNATURES table;
_configure_function_signatures calls getattr(lib, name) for each) and
python/quantkernel/_abi.py:3-4 (ABI_MAJOR, ABI_MINOR).  Run where the
reference sources are present:

    python tests/golden/gen_loader_symbols.py /root/reference
"""
import json
import re
import sys
from pathlib import Path


def main(ref_root):
    ref = Path(ref_root)
    loader = (ref / "python" / "quantkernel" / "_loader.py").read_text()
    table = loader.split("_FUNCTION_SIGNATURES = [", 1)[1].split("\n]\n", 1)[0]
    names = re.findall(r'^\s*\("(qk_\w+)"', table, re.M)
    abi = (ref / "python" / "quantkernel" / "_abi.py").read_text()
    major = int(re.search(r"^ABI_MAJOR\s*=\s*(\d+)", abi, re.M).group(1))
    minor = int(re.search(r"^ABI_MINOR\s*=\s*(\d+)", abi, re.M).group(1))
    out = Path(__file__).with_name("loader_symbols.json")
    out.write_text(json.dumps({"abi_major": major, "abi_minor": minor, "symbols": names},
                              indent=1) + "\n")
    print(f"{len(names)} symbols, ABI {major}.{minor} -> {out.name}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
