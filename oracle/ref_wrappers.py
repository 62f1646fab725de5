"""Writes _ref/<config>/bodies.inc: one C++ function per reference kernel
body (operator family x {dslash, dslash_achimbdpsi} x isign x 12/18 links
x boundary conditions simple in every dimension, "ssss", or twisted in every
dimension, "tttt", in x y z t order) that binds the parameter names the
generated body uses (codegen/jinja/*_decl.h.j2) from a RefArgs struct and
then #includes the body the reference's generator wrote (file names:
codegen/dslash.cc:617-630).

TEST INFRASTRUCTURE ONLY (part of the oracle/_ref build).

usage: python ref_wrappers.py OUT.inc FTNAME VECLEN SOALEN
"""
import sys

FAMILIES = [("", "wilson", "void"), ("clov_{ft}_", "clover", "CloverBlock"),
            ("tm_", "tm", "void"), ("tm_clov_{ft}_", "tmclover", "FullCloverBlock")]


def main(out, ft, v, s):
    lines = ["// generated by oracle/ref_wrappers.py -- do not edit", ""]
    tables = {}
    for bc in ("ssss", "tttt"):
        table = tables.setdefault(bc, [])
        for prefix, fam, clt in FAMILIES:
            for kernel in ("dslash", "dslash_achimbdpsi"):
                for pm in ("plus", "minus"):
                    for recon in (12, 18):
                        c12 = "true" if recon == 12 else "false"
                        name = f"ref_{fam}_{kernel}_{pm}_{recon}_{bc}"
                        body = (f"generated/{prefix.format(ft=ft)}{kernel}_{pm}_body_{ft}_{ft}"
                                f"_v{v}_s{s}_{recon}_{bc}")
                        cl = "" if clt == "void" else (
                            f"  const Ty::{clt} *clBase = (const Ty::{clt} *)A.clBase;\n")
                        lines.append(f"static void {name}(const RefArgs &A)\n{{\n"
                                     f"  typedef QPhiX::Types<FT, VECLEN, SOALEN, {c12}> Ty;\n"
                                     f"  REF_BIND_COMMON\n{cl}"
                                     f"#include \"{body}\"\n}}\n")
                        table.append(name)
    for bc, tname in (("ssss", "ref_table"), ("tttt", "ref_table_tbc")):
        table = tables[bc]
        lines.append("// [family][achimbdpsi][minus][recon18]")
        lines.append(f"static const ref_body_fn {tname}[4][2][2][2] = {{")
        k = 0
        for f in range(4):
            lines.append("  {")
            for a in range(2):
                row = []
                for m in range(2):
                    row.append("{" + ", ".join(table[k + r] for r in range(2)) + "}")
                    k += 2
                lines.append("    {" + ", ".join(row) + "},")
            lines.append("  },")
        lines.append("};")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
