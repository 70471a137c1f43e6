#!/usr/bin/env python3
"""hipgen.py -- HIP emitter for sincFeyn's generated integrand kernels (gfx950).

Input: one generated kernel file of the reference, kernels/<class>/<graph>.c.  Each holds the integrand
twice, once written against REAL = double (with SIMD helpers for the host) and once against REAL = QREAL,
the high-precision type (reference driver/common.h:5-22, kernels/g2/g2_00.c:231-460).  The QREAL half is
a complete scalar copy of the integrand: chunk functions bk_/ab_ (the forest blocks of large kernels), the
point function pt_, helpers vp_/p_/nd_ and the Sinc wrapper f_, all plain + - * / on REAL.  The emitter
turns that half into ONE set of C++ function templates over REAL, which the runtime instantiates twice on
the device: REAL = double for the bulk of the points and REAL = sfx_dd for the fallback.  So both device
paths come from the same text, and the double path is the arithmetic of the reference's own GPU build
(which compiles the same half as double, reference tools/gpugen.py:20-29).

Transforms (text level; the kernels are regular machine-written C):
  * `static REAL FN(name)(params){`    -> `template<class REAL> SFX_FN REAL name(params, const int SQ_AB){`
  * `static void FN(bk_x)(const REAL *F, REAL *T){` (chunk) -> the body `bk_x_b`, an out-of-line copy
    `bk_x_c` (SFX_CHUNK) and a dispatcher `bk_x` that inlines the body where SFX_CHUNK_INLINE(REAL) holds
    (the FP64 kernels: no call frame, F and T stay in registers) and calls the copy otherwise (the
    double-double kernels, whose inlined bodies would take the compiler many minutes)
  * every call `FN(name)(args)`         -> `name<REAL>(args, SQ_AB)`  (chunk calls get no SQ_AB)
  * the Gauss-Legendre tables at file top -> `__constant__` arrays
  * a point function whose forest nodes are all leaves gets a second body for the FP64 kernels, a loop
    with one switch case per node (node_split; the register need of one forest block instead of all)
  * SQ_AB (the reference's thread-local "compute the |.| sums" flag, common.h:24-26) becomes a parameter;
    the runtime passes a compile-time constant, so with 0 the companion sums are dead code.
The graph descriptor line `const graph_t G_<name>={...}` (common.h:38) gives name, Volkov string, loops L,
dimension D and mirror weight.  The emitter also counts the FP64 + - * / of every function body (a
static count, loops expanded by their literal bounds where the code has them), for the roofline report.

    python3 -m sincfeyn_amd.emit.hipgen KERNEL.c OUT_DIR     -> OUT_DIR/k_<name>.hip and k_<name>.json
"""
import json
import os
import re
import sys

DESC_RE = re.compile(r'^const graph_t G_(\w+)=\{"([^"]*)","([^"]*)",(\d+),(\d+),(\d+),(\d+),(\d+),([^}]*)\}', re.M)
GL_RE = re.compile(r'^static const double (GL[A-Z]*X_\w+)\[(\d+)\]=\{([^}]*)\}, (GL[A-Z]*W_\w+)\[(\d+)\]=\{([^}]*)\};', re.M)
DEF_RE = re.compile(r'^static (REAL|void) FN\((\w+)\)\(', re.M)


class EmitError(Exception):
    pass


def _match_paren(s, i):
    """index of the ')' closing the '(' at s[i]"""
    depth = 0
    for p in range(i, len(s)):
        c = s[p]
        if c == '(':
            depth += 1
        elif c == ')':
            depth -= 1
            if depth == 0:
                return p
    raise EmitError('unbalanced parentheses')


def parse_kernel(text):
    """(descriptor dict, GL tables, QREAL half text) of one generated kernel file"""
    m = DESC_RE.search(text)
    if not m:
        raise EmitError('no graph_t descriptor')
    if m.group(2) != m.group(1):
        raise EmitError('descriptor G_%s names graph %s' % (m.group(1), m.group(2)))
    # the point function (ptq, common.h:38) is present for graphs with a single-point body; VP graphs
    # (vacuum-polarisation insertions) have 0 there
    fns = [t.strip() for t in m.group(9).split(',')]
    desc = dict(name=m.group(1), volkov=m.group(3), L=int(m.group(4)), D=int(m.group(5)),
                weight=int(m.group(6)), n=int(m.group(7)), fixed=int(m.group(8)),
                has_pt=len(fns) >= 4 and fns[3] == 'pt_%s_q' % m.group(1))
    start = text.find('#define REAL QREAL')
    if start < 0:
        raise EmitError('no QREAL half')
    fn = text.find('#define FN(x) x##_q', start)
    if fn < 0:
        raise EmitError('no FN(x) x##_q after #define REAL QREAL')
    body_start = fn + len('#define FN(x) x##_q')
    end = text.find('#undef REAL', body_start)
    if end < 0:
        raise EmitError('QREAL half not closed by #undef REAL')
    tables = [(g[0], int(g[1]), g[2], g[3], int(g[4]), g[5]) for g in GL_RE.findall(text)]
    return desc, tables, text[body_start:end]


def _is_chunk(name):
    return name.startswith('bk_') or name.startswith('ab_')


def transform(body):
    """QREAL half -> templated device code; returns (code, function names)"""
    defs = DEF_RE.findall(body)
    if not defs:
        raise EmitError('no functions in the QREAL half')
    names = [n for _, n in defs]
    chunk = {n for n in names if _is_chunk(n)}
    # calls first: every FN(name)( not preceded by "static REAL " / "static void "
    out, pos = [], 0
    for m in re.finditer(r'FN\((\w+)\)\(', body):
        name = m.group(1)
        if name not in names:
            raise EmitError('call of an unknown function %s' % name)
        head = body[max(0, m.start() - 12):m.start()]
        if re.search(r'static (REAL|void) $', head):
            continue
        close = _match_paren(body, m.end() - 1)
        out.append(body[pos:m.start()])
        out.append('%s<REAL>(' % name)
        args = body[m.end():close]
        out.append(args)
        if name not in chunk:
            out.append(', SQ_AB')
        pos = close
    out.append(body[pos:])
    body = ''.join(out)
    # definitions
    res, pos = [], 0
    for m in re.finditer(r'^static (REAL|void) FN\((\w+)\)\(', body, re.M):
        ret, name = m.group(1), m.group(2)
        close = _match_paren(body, m.end() - 1)
        params = body[m.end():close]
        res.append(body[pos:m.start()])
        if name in chunk:
            if params.strip() != 'const REAL *F, REAL *T':
                raise EmitError('chunk %s has parameters (%s), not (const REAL *F, REAL *T)' % (name, params))
            pr = 'const REAL *__restrict__ F, REAL *__restrict__ T'
            # the body (always inlined into its caller), an out-of-line copy, and the dispatcher the point
            # function calls: inline where SFX_CHUNK_INLINE(REAL) holds (FP64), else the out-of-line call
            res.append('template<class REAL> SFX_FN void %s_b(%s);\n' % (name, pr))
            res.append('template<class REAL> SFX_CHUNK void %s_c(%s){ %s_b<REAL>(F, T); }\n' % (name, pr, name))
            res.append('template<class REAL> SFX_FN void %s(%s){\n'
                       '  if constexpr(SFX_CHUNK_INLINE(REAL)) %s_b<REAL>(F, T); else %s_c<REAL>(F, T); }\n'
                       % (name, pr, name, name))
            res.append('template<class REAL> SFX_FN %s %s_b(%s' % (ret, name, pr))
        else:
            res.append('template<class REAL> SFX_FN %s %s(%s, const int SQ_AB' % (ret, name, params))
        pos = close
    res.append(body[pos:])
    code = ''.join(res)
    if re.search(r'\bstatic\b', code):
        raise EmitError('unexpected static declaration left in the QREAL half')
    if re.search(r'\bFN\(', code):
        raise EmitError('FN( left after the transform')
    return code, names


_PT_HEAD = 'template<class REAL> SFX_FN REAL pt_%s(const REAL *x, const REAL Vx, REAL *ab, const int SQ_AB){'
_XLOAD_RE = re.compile(r'^  const REAL (x\d+)=x\[\d+\];$')
_FINAL_RE = re.compile(r'^  \{ const REAL iV=1/\((n_\w+)_V\+Vx\);')


def node_split(code, name, groups=None):
    """The point function once more as a loop over its forest nodes, for the FP64 kernels (see
    SFX_NODE_SPLIT in sfx_integrand.hpp): one switch case per node, or per run of consecutive small nodes
    (forest block under half the largest, so they keep their common subexpressions), each case followed
    by that node's term of the final sum.  Applies when every node is a leaf and the final sum takes one
    term per node in node order (the terms then run in the same order as in the straight-line form);
    returns (code, number of cases), or (code unchanged, 0) when the point function has another shape.
    groups (lists of consecutive node indices, in order) replaces the grouping rule; for g3_lbl, the rule's
    six large nodes alone and the six small ones together measured best (244 VGPRs; two large nodes in
    one case need more than 256 and drop to one wave per SIMD)."""
    lines = code.split('\n')
    head = _PT_HEAD % name
    if head not in lines:
        return code, 0
    s = lines.index(head)
    e = lines.index('}', s)
    body = lines[s + 1:e]
    nodes = [i for i, l in enumerate(body) if l.startswith('  /* node')]
    if len(nodes) < 2 or any('kids ()' not in body[i] for i in nodes):
        return code, 0
    pre = body[:nodes[0]]
    xs = [_XLOAD_RE.match(l) for l in pre]
    if not pre or not all(xs):
        return code, 0
    xs = [m.group(1) for m in xs]
    try:
        t = body.index('  REAL tot=0, tota=0;')
    except ValueError:
        return code, 0
    secs = [body[a:b] for a, b in zip(nodes, nodes[1:] + [t])]
    fin, i, rest = [], 0, body[t + 1:]
    while i < len(rest) and _FINAL_RE.match(rest[i]):
        j = i
        while not rest[j].endswith('tota+=ra; }'):
            j += 1
        fin.append((_FINAL_RE.match(rest[i]).group(1), rest[i:j + 1]))
        i = j + 1
    if rest[i:] != ['  *ab=tota; return tot;'] or len(fin) != len(secs):
        return code, 0
    cost = []
    for sec, (nm, _) in zip(secs, fin):
        if not sec[1].startswith('  REAL %s_W[' % nm):
            return code, 0
        calls = re.findall(r'\b(bk_\w+)<REAL>\(F_, T_\)', '\n'.join(sec))
        c = len(sec)
        for ch in calls:
            hb = 'template<class REAL> SFX_FN void %s_b(const REAL *__restrict__ F, REAL *__restrict__ T){' % ch
            if hb in lines:
                b0 = lines.index(hb)
                c += lines.index('}', b0) - b0
        cost.append(c)
    big = max(cost)
    if groups is not None:
        return _node_cases(lines, s, e, body, pre, xs, secs, fin, groups)
    groups = []
    for k, c in enumerate(cost):
        if 2*c >= big or not groups or 2*cost[groups[-1][-1]] >= big:
            groups.append([k])
        else:
            groups[-1].append(k)
    return _node_cases(lines, s, e, body, pre, xs, secs, fin, groups)


def _node_cases(lines, s, e, body, pre, xs, secs, fin, groups):
    """the point function's text with its second, switch-case body (node_split)"""
    xre = re.compile(r'\b(%s)\b' % '|'.join(xs))
    sw = ['  } else {'] + pre + ['  REAL tot=0, tota=0;',
          '  #pragma clang loop unroll(disable)',
          '  for(int k_=0; k_<%d; k_++){ int kk_=k_; sfx_opaque_s(kk_); switch(kk_){' % len(groups)]
    for gi, g in enumerate(groups):
        sw += ['  case %d: {' % gi, '  ' + ' '.join('REAL q%s=%s; sfx_opaque(q%s);' % (x, x, x) for x in xs)]
        for k in g:
            sw += [xre.sub(lambda m: 'q' + m.group(1), l) for l in secs[k]] + fin[k][1]
        sw += ['  } break;']
    sw += ['  default: break;', '  } }', '  *ab=tota; return tot;', '  }']
    out = lines[:s + 1] + ['  if constexpr(!SFX_NODE_SPLIT(REAL)){'] + body + sw + lines[e:]
    return '\n'.join(out), len(groups)


_OPS_RE = re.compile(r'(?<![eE])[-+*/](?![=+\-])')


def count_ops(body):
    """static count of the binary + - * / in the statements of the QREAL half (compound assignments
    count once; unary minus after '(' '=' ',' is not counted).  Informational: the roofline uses the
    per-point counts measured by the tests, this is the emitter's own upper estimate."""
    n = 0
    for line in body.splitlines():
        s = line.strip()
        if not s or s.startswith('/*') or s.startswith('#') or s.startswith('for('):
            continue
        s = re.sub(r'\[[^\]]*\]', '[]', s)
        s = re.sub(r'(^|[=(,?:])\s*-', r'\1', s)
        n += len(_OPS_RE.findall(s)) + len(re.findall(r'[-+*/]=', s))
    return n


def emit(src_path, out_dir):
    with open(src_path) as f:
        text = f.read()
    desc, tables, body = parse_kernel(text)
    code, names = transform(body)
    name = desc['name']
    if 'f_' + name not in names:
        raise EmitError('%s: no f_%s in the QREAL half' % (src_path, name))
    code, ncases = node_split(code, name)
    inc = ['// generated by sincfeyn_amd/emit/hipgen.py from %s -- do not edit' % os.path.basename(src_path),
           '// the integrand of graph %s as function templates over REAL' % name]
    for xn, xl, xv, wn, wl, wv in tables:
        inc.append('static __constant__ const double %s[%d] = {%s}, %s[%d] = {%s};' % (xn, xl, xv, wn, wl, wv))
    inc += ['#define RABS(x) sfx_abs(x)', code.strip('\n'), '#undef RABS', '']
    lines = ['// generated by sincfeyn_amd/emit/hipgen.py from %s -- do not edit' % os.path.basename(src_path),
             '// graph %s: %s, L=%d, D=%d, weight %d' % (name, desc['volkov'], desc['L'], desc['D'], desc['weight']),
             '#include "sfx_integrand.hpp"',
             'namespace sfxk_%s {' % name,
             '#include "k_%s.inc"' % name,
             'struct Graph {',
             '    static constexpr int D = %d;' % desc['D'],
             '    template<class REAL> static __device__ __forceinline__ REAL f(const REAL *sf, REAL *ab, const int sqab)',
             '    { return f_%s<REAL>(sf, ab, sqab); }' % name,
             '    static constexpr int NL = %d, FIXED = %d;' % (desc['n'], desc['fixed']),
             '    static constexpr bool HAS_PT = %s;' % ('true' if desc['has_pt'] else 'false'),
             '    static constexpr bool QD = %s;     // quad-double recompute (two-loop graphs)' % (
                 'true' if desc['L'] == 2 else 'false'),
             '    template<class REAL> static __device__ __forceinline__ REAL pt(const REAL *x, REAL *ab)',
             ('    { return pt_%s<REAL>(x, (REAL)0, ab, 1); }' % name) if desc['has_pt'] else
             '    { (void)x; *ab = (REAL)0; return (REAL)0; }',
             '};',
             '}  // namespace sfxk_%s' % name,
             'SFX_GRAPH_ENTRY(%s, sfxk_%s::Graph, "%s", %d, %d, %d)' % (
                 name, name, desc['volkov'].replace('"', '\\"'), desc['L'], desc['D'], desc['weight']),
             '']
    os.makedirs(out_dir, exist_ok=True)
    for ext, body_lines in (('inc', inc), ('hip', lines)):
        out = os.path.join(out_dir, 'k_%s.%s' % (name, ext))
        text = '\n'.join(body_lines)
        if os.path.exists(out) and open(out).read() == text:
            continue                      # unchanged: keep the timestamp, so an incremental build skips it
        with open(out + '.tmp', 'w') as f:
            f.write(text)
        os.replace(out + '.tmp', out)
    out = os.path.join(out_dir, 'k_%s.hip' % name)
    meta = dict(desc, source=os.path.basename(src_path), functions=len(names),
                chunk_functions=sum(1 for n in names if _is_chunk(n)), static_ops=count_ops(body),
                node_split=ncases)
    mp = os.path.join(out_dir, 'k_%s.json' % name)
    text = json.dumps(meta, indent=1)
    if not (os.path.exists(mp) and open(mp).read() == text):
        with open(mp + '.tmp', 'w') as f:     # replaced whole: a build running beside this one reads it
            f.write(text)
        os.replace(mp + '.tmp', mp)
    return out, meta


def main(argv):
    if len(argv) != 3:
        print(__doc__)
        return 2
    out, meta = emit(argv[1], argv[2])
    print(out, meta['name'], 'D=%d' % meta['D'])
    return 0


if __name__ == '__main__':
    sys.exit(main(sys.argv))
