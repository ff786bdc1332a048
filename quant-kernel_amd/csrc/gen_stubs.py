"""Generates the loader-compatibility stubs of libquantkernel.so.

The reference's Python loader (python/quantkernel/_loader.py:91-316) resolves
every function of its qk_api.h when it loads the library.  This library
implements the Monte Carlo path only; for each other exported function of the
reference header this script writes

  * a declaration into include/quantkernel/qk_api_stubs.h (same name, return
    type and parameter types as the reference's declaration), and
  * a definition into quant-kernel_amd/csrc/qk_stubs.cpp that computes
    nothing: scalar functions return NaN and batch functions return
    QK_ERR_RUNTIME_INIT, both with "<name>: not provided ..." as the last
    error, so the reference wrapper raises instead of returning numbers.

Run it where the reference sources are present (it is not part of the build):

    python quant-kernel_amd/csrc/gen_stubs.py /root/reference/cpp/include/quantkernel/qk_api.h
"""
import re
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
OWN = ROOT / "include" / "quantkernel" / "qk_api.h"
HDR_OUT = ROOT / "include" / "quantkernel" / "qk_api_stubs.h"
SRC_OUT = ROOT / "quant-kernel_amd" / "csrc" / "qk_stubs.cpp"

DECL = re.compile(r"QK_EXPORT\s+([\w\s\*]+?)\s+(qk_\w+)\s*\(([^;]*?)\)\s*;", re.S)


def declarations(text):
    out = []
    for ret, name, params in DECL.findall(text):
        ps = [" ".join(p.split()) for p in params.split(",") if p.strip()]
        out.append((" ".join(ret.split()), name, ps))
    return out


def param_type(p):
    # drop the parameter name: the last identifier of "const double* spot"
    m = re.match(r"^(.*?[\s\*])(\w+)$", p)
    return " ".join(m.group(1).split()).replace(" *", "*") if m else p


def main(ref_header):
    ref = declarations(Path(ref_header).read_text())
    own = {name for _, name, _ in declarations(OWN.read_text())}
    stubs = [d for d in ref if d[1] not in own]
    for ret, name, _ in stubs:
        if ret not in ("double", "int32_t"):
            raise SystemExit(f"unexpected return type {ret!r} of {name}")
    h = ["/* qk_api_stubs.h - loader-compatibility stubs (generated by",
         " * quant-kernel_amd/csrc/gen_stubs.py; do not edit).",
         " *",
         " * The functions of the reference's qk_api.h outside the Monte Carlo path,",
         " * with the reference's names and parameter lists, so that its Python loader",
         " * (python/quantkernel/_loader.py:91-316), which resolves every symbol when it",
         " * loads, accepts this library unchanged.  None of them computes anything:",
         " * scalar functions return NaN and batch functions return QK_ERR_RUNTIME_INIT,",
         " * each with \"<name>: not provided ...\" as the last error. */",
         "#ifndef QK_API_STUBS_H", "#define QK_API_STUBS_H", "", '#include "qk_api.h"', "",
         "#ifdef __cplusplus", 'extern "C" {', "#endif", ""]
    for ret, name, ps in stubs:
        h.append(f"QK_EXPORT {ret} {name}({', '.join(ps) if ps else 'void'});")
    h += ["", "#ifdef __cplusplus", "}", "#endif", "", "#endif /* QK_API_STUBS_H */", ""]
    c = ["// qk_stubs.cpp - loader-compatibility stubs (generated by gen_stubs.py; do",
         "// not edit): the reference qk_api.h functions outside the Monte Carlo path.",
         "// Scalars return NaN, batches QK_ERR_RUNTIME_INIT, each with a last error.",
         '#include "quantkernel/qk_api_stubs.h"', "",
         "namespace qk {", "void set_last_error(const char* msg);", "}", "",
         "namespace {", "",
         "double stub_nan(const char* msg) {", "    qk::set_last_error(msg);",
         "    return __builtin_nan(\"\");", "}", "",
         "int32_t stub_rc(const char* msg) {", "    qk::set_last_error(msg);",
         "    return QK_ERR_RUNTIME_INIT;", "}", "", "}  // namespace", "", 'extern "C" {', ""]
    for ret, name, ps in stubs:
        msg = f"{name}: not provided by the MI355X Monte Carlo library"
        types = ", ".join(param_type(p) for p in ps) if ps else "void"
        call = "stub_nan" if ret == "double" else "stub_rc"
        c.append(f'{ret} {name}({types}) {{ return {call}("{msg}"); }}')
    c += ["", '}  // extern "C"', ""]
    HDR_OUT.write_text("\n".join(h))
    SRC_OUT.write_text("\n".join(c))
    print(f"{len(stubs)} stubs ({sum(r == 'double' for r, _, _ in stubs)} scalar) -> {HDR_OUT.name}, {SRC_OUT.name}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/cpp/include/quantkernel/qk_api.h")
