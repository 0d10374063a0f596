"""TEST INFRASTRUCTURE: the op-counting build of the CPU restatement.

Rewrites the LLVM IR of oracle/mop_oracle.c (clang -O2 without inlining,
vectorisation or FMA contraction) so that every basic block, when it runs,
adds one to a counter indexed by (phase, block).  The floating-point work of
each block is tallied statically here, by kind, and written to a JSON table
beside the library; the frame's operation count is then

    ops[phase][kind] = sum over blocks b of runs[phase][b] * static[b][kind]

which is exact for every block that runs to its end (C code has no other
kind).  A phase is set on entry to the functions of PHASE_OF (and restored on
return); every other function counts in its caller's phase.

Counted as one operation per lane: fadd/fsub ("add"), fmul ("mul"),
fdiv/frem ("div"), fneg ("neg"), fcmp ("cmp"), float<->int and
float<->double conversions ("cvt"), selects of floating-point values
("select"), and the calls sqrt, min/max, fabs, floor/ceil/round/trunc and the
C library's transcendental functions (each under its own name).  Operations on
double are kept apart with an "f64_" prefix.  Integer and address arithmetic
is not counted.

Usage: python3 opcount.py IN.ll OUT.ll OUT.json
"""
import json
import re
import sys

PHASES = ["other", "vertex", "triangle", "walk", "shade", "ordered", "shadow", "post",
          "background", "skin", "clear"]

# Functions of oracle/mop_oracle.c that begin a phase (the stage each stands
# for in the device pipeline: k_vertex, k_tri/k_clip setup, the raster walks,
# k_shade, k_tile_ordered, the shadow kernels, k_tail and the post kernels,
# k_background, k_skin/k_morph, the clear).
PHASE_OF = {
    "xform_vertex": "vertex",
    "draw_one": "triangle",
    "walk_vis": "walk",
    "walk_vis_flat": "walk",
    "shade_segment": "shade",
    "draw_ordered": "ordered",
    "mop_oracle_shadow_render_mesh": "shadow",
    "resolve": "post",
    "mop_oracle_postprocess_apply": "post",
    "mop_oracle_draw_outline": "post",
    "mop_oracle_framebuffer_read_color_downsampled": "post",
    "mop_oracle_draw_background_hdri": "background",
    "mop_oracle_skin_apply": "skin",
    "mop_oracle_morph_apply": "skin",
    "mop_oracle_frame_begin": "clear",
}

ARITH = {"fadd": "add", "fsub": "add", "fmul": "mul", "fdiv": "div", "frem": "div",
         "fneg": "neg", "fcmp": "cmp", "fptosi": "cvt", "fptoui": "cvt", "sitofp": "cvt",
         "uitofp": "cvt", "fpext": "cvt", "fptrunc": "cvt"}
INTRINSIC = [("llvm.sqrt.", "sqrt"), ("llvm.minnum.", "minmax"), ("llvm.maxnum.", "minmax"),
             ("llvm.minimum.", "minmax"), ("llvm.maximum.", "minmax"), ("llvm.fabs.", "abs"),
             ("llvm.floor.", "round"), ("llvm.ceil.", "round"), ("llvm.trunc.", "round"),
             ("llvm.round.", "round"), ("llvm.rint.", "round"), ("llvm.nearbyint.", "round"),
             ("llvm.fma.", "fma"), ("llvm.fmuladd.", "fma"), ("llvm.copysign.", "abs")]
LIBM = {"sqrtf": "sqrt", "sqrt": "sqrt", "fabsf": "abs", "fabs": "abs", "floorf": "round",
        "ceilf": "round", "fminf": "minmax", "fmaxf": "minmax"}
LIBM_NAMED = {"powf", "pow", "expf", "exp", "exp2f", "logf", "log", "log2f", "sinf", "sin",
              "cosf", "cos", "tanf", "tan", "atanf", "atan", "atan2f", "atan2", "asinf",
              "asin", "acosf", "acos", "fmodf", "fmod", "sinhf", "coshf", "tanhf", "cbrtf"}

LABEL = re.compile(r'^([A-Za-z$._0-9-]+|"[^"]+"):')
DEFINE = re.compile(r'^define [^@]*@("[^"]+"|[A-Za-z$._0-9-]+)\(')
VEC = re.compile(r'<(\d+) x (float|double|half)>')
CALL = re.compile(r'\bcall\b[^@]*@([A-Za-z$._0-9-]+)\(')


def lanes_and_prec(text):
    """(lanes, is_double) of the first floating-point type in text."""
    m = VEC.search(text)
    fm = re.search(r'\b(float|double|half)\b', text)
    if m and (not fm or m.start() <= fm.start()):
        return int(m.group(1)), m.group(2) == "double"
    if fm:
        return 1, fm.group(1) == "double"
    return 1, False


def classify(line):
    """{kind: count} of one IR instruction line."""
    s = line.strip()
    if not s or s.startswith(";"):
        return {}
    rhs = s.split("=", 1)[1].strip() if re.match(r'^%[^ ]+ = ', s) else s
    op = rhs.split(" ", 1)[0]
    if op in ARITH:
        rest = rhs[len(op):]
        if op == "fcmp":
            rest = rest.strip().split(" ", 1)[1] if " " in rest.strip() else rest
        n, dbl = lanes_and_prec(rest)
        if op in ("sitofp", "uitofp"):
            # the lane count comes from the source, the precision from the result
            src = rest.split(" to ")[0]
            dst = rest.split(" to ")[1] if " to " in rest else rest
            n = lanes_and_prec(src)[0] if VEC.search(src) else 1
            dbl = lanes_and_prec(dst)[1]
        return {("f64_" if dbl else "") + ARITH[op]: n}
    if op == "select":
        # select i1 %c, <ty> %a, <ty> %b: count floating-point selects only
        parts = rhs.split(",", 2)
        if len(parts) >= 2 and re.search(r'\b(float|double|half)\b', parts[1]):
            n, dbl = lanes_and_prec(parts[1])
            return {("f64_" if dbl else "") + "select": n}
        return {}
    m = CALL.search(rhs)
    if m and (op in ("call", "tail", "musttail", "notail")):
        name = m.group(1)
        for pre, kind in INTRINSIC:
            if name.startswith(pre):
                suf = name[len(pre):]  # f32, f64, v2f32, ...
                vm = re.match(r'v(\d+)', suf)
                return {("f64_" if "f64" in suf else "") + kind: int(vm.group(1)) if vm else 1}
        if name in LIBM:
            return {LIBM[name]: 1}
        if name in LIBM_NAMED:
            return {"libm_" + name: 1}
    return {}


def instrument(src_lines):
    out, blocks = [], []
    nph = len(PHASES)
    i, n = 0, len(src_lines)
    while i < n:
        line = src_lines[i]
        m = DEFINE.match(line)
        if not m or not line.rstrip().endswith("{"):
            out.append(line)
            i += 1
            continue
        fname = m.group(1).strip('"')
        phase = PHASE_OF.get(fname)
        out.append(line)
        i += 1
        # collect the body up to the closing brace
        body = []
        while not src_lines[i].startswith("}"):
            body.append(src_lines[i])
            i += 1
        close = src_lines[i]
        i += 1
        # split into blocks: (label line or None, [instruction lines])
        blks, cur_label, cur = [], None, []
        for b in body:
            if LABEL.match(b):
                blks.append((cur_label, cur))
                cur_label, cur = b, []
            else:
                cur.append(b)
        blks.append((cur_label, cur))
        for bi, (lab, ins) in enumerate(blks):
            if lab is not None:
                out.append(lab)
            # phis (and blank/comment lines before them) stay first
            k = 0
            while k < len(ins) and (not ins[k].strip() or ins[k].strip().startswith(";")
                                    or " = phi " in ins[k]):
                out.append(ins[k])
                k += 1
            bid = len(blocks)
            counts = {}
            for x in ins:
                for kind, c in classify(x).items():
                    counts[kind] = counts.get(kind, 0) + c
            blocks.append({"fn": fname, "ops": counts})
            if bi == 0 and phase is not None:
                out.append(f"  %__opc_saved = call i32 @mop_opc_enter(i32 {PHASES.index(phase)})")
            out.append(f"  call void @mop_opc_bb(i32 {bid})")
            for x in ins[k:]:
                if phase is not None and re.match(r'^\s+ret\b', x):
                    out.append("  call void @mop_opc_leave(i32 %__opc_saved)")
                out.append(x)
        out.append(close)
    # memory attributes would let the code generator treat the counting calls
    # as removable: drop them
    res = []
    for line in out:
        if line.startswith("attributes #"):
            line = re.sub(r'\bmemory\([^)]*\)', "", line)
            line = re.sub(r'\b(readnone|readonly|writeonly)\b', "", line)
        res.append(line)
    res.append("")
    res.append("declare void @mop_opc_bb(i32)")
    res.append("declare i32 @mop_opc_enter(i32)")
    res.append("declare void @mop_opc_leave(i32)")
    res.append(f"@mop_opc_nblocks = constant i32 {len(blocks)}")
    res.append(f"@mop_opc_counts = global [{nph * len(blocks)} x i64] zeroinitializer")
    return res, blocks


def main():
    src, dst, table = sys.argv[1:4]
    with open(src) as f:
        lines = f.read().split("\n")
    res, blocks = instrument(lines)
    with open(dst, "w") as f:
        f.write("\n".join(res) + "\n")
    with open(table, "w") as f:
        json.dump({"phases": PHASES, "blocks": blocks}, f)


if __name__ == "__main__":
    main()
