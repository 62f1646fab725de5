"""Writes _ref/<config>/bodies.inc: one C++ function per reference kernel
body (operator family x {dslash, dslash_achimbdpsi} x isign x 12/18 links
x boundary conditions simple in every dimension, "ssss", or twisted in every
dimension, "tttt", in x y z t order) that binds the parameter names the
generated body uses (codegen/jinja/*_decl.h.j2) from a RefArgs struct and
then #includes the body the reference's generator wrote (file names:
codegen/dslash.cc:617-630).

With --faces, writes _ref/<config>/faces.inc instead: the face bodies.
face_proj_dir_{plus,minus} (codegen/jinja/dslash_decl.h.j2:90-99) as one
function per isign switching on dir over the eight
dslash_face_pack_to_<back|forw>_<X|Y|Z|T> bodies, and
face_finish_dir_*, face_clov_finish_dir_*, tm_face_finish_dir_*,
tm_clov_face_finish_dir_* (codegen/jinja/*_dslash_decl.h.j2:90-115) as one
function per family x isign x 12/18 links x simple|twisted boundary,
switching on dir over the eight *dslash_face_unpack_from_* bodies (the
dispatch of codegen/jinja/dslash_spec.h.j2:1-80; file names
codegen/dslash.cc:664-744).  Parameters are bound from a FaceArgs struct.

TEST INFRASTRUCTURE ONLY (part of the oracle/_ref build).

usage: python ref_wrappers.py [--faces] OUT.inc FTNAME VECLEN SOALEN
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


DIRS = [(fb, d) for d in "XYZT" for fb in ("back", "forw")]  # dir = 2 dim + fb


def _switch(bodies):
    """A switch on dir with body k #included in case k."""
    out = ["  switch (F.dir) {"]
    for k, b in enumerate(bodies):
        out.append(f"  case {k}: {{\n#include \"{b}\"\n  }} break;")
    out.append("  default: break;\n  }")
    return "\n".join(out)


def faces(out, ft, v, s):
    lines = ["// generated by oracle/ref_wrappers.py --faces -- do not edit", ""]
    pack = []
    for pm in ("plus", "minus"):
        name = f"ref_face_pack_{pm}"
        bodies = [f"generated/dslash_face_pack_to_{fb}_{d}_{pm}_{ft}_{ft}_v{v}_s{s}" for fb, d in DIRS]
        lines.append(f"static void {name}(const FaceArgs &F)\n{{\n"
                     f"  typedef QPhiX::Types<FT, VECLEN, SOALEN, true> Ty;\n"
                     f"  REF_BIND_FACE\n{_switch(bodies)}\n}}\n")
        pack.append(name)
    lines.append(f"static const ref_face_fn ref_face_pack_table[2] = {{{', '.join(pack)}}};")
    tables = {}
    for bc in ("s", "t"):
        table = tables.setdefault(bc, [])
        for prefix, fam, clt in FAMILIES:
            for pm in ("plus", "minus"):
                for recon in (12, 18):
                    c12 = "true" if recon == 12 else "false"
                    name = f"ref_{fam}_face_finish_{pm}_{recon}_{bc}"
                    bodies = [f"generated/{prefix.format(ft=ft)}dslash_face_unpack_from_{fb}_{d}_{pm}_{ft}_{ft}"
                              f"_v{v}_s{s}_{recon}_{bc}" for fb, d in DIRS]
                    cl = "" if clt == "void" else (
                        f"  const Ty::{clt} *clBase = (const Ty::{clt} *)F.clBase;\n")
                    lines.append(f"static void {name}(const FaceArgs &F)\n{{\n"
                                 f"  typedef QPhiX::Types<FT, VECLEN, SOALEN, {c12}> Ty;\n"
                                 f"  REF_BIND_FACE\n{cl}{_switch(bodies)}\n}}\n")
                    table.append(name)
    for bc, tname in (("s", "ref_face_table"), ("t", "ref_face_table_tbc")):
        t = tables[bc]
        lines.append("// [family][minus][recon18]")
        lines.append(f"static const ref_face_fn {tname}[4][2][2] = {{")
        for f in range(4):
            rows = ["{" + ", ".join(t[f * 4 + m * 2 + r] for r in range(2)) + "}" for m in range(2)]
            lines.append("  {" + ", ".join(rows) + "},")
        lines.append("};")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    if sys.argv[1] == "--faces":
        faces(sys.argv[2], sys.argv[3], int(sys.argv[4]), int(sys.argv[5]))
    else:
        main(sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4]))
