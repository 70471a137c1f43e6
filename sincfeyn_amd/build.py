"""build.py -- builds the MI355X library sincfeyn_amd/_lib/libsfx.so and the host CLI _lib/sincqed.

  1. the HIP emitter (emit/hipgen.py) turns every generated integrand kernel of the selected classes,
     read from the sincFeyn tree (SFX_KERNELS, default /root/reference/kernels), into _gen/k_<graph>.hip;
  2. _gen/sfx_registry.cpp lists the graphs;
  3. hipcc compiles every translation unit for gfx950 (FP contraction off, packed-FP32 ops off) in
     parallel, and links the shared library;
  4. the C host layer src/sincqed.c is linked against it.

Where the kernel sources are absent (a machine that only runs the library), the emitted sources already
in _gen/ are compiled as they are.

    python3 -m sincfeyn_amd.build [--force] [--jobs N] [--classes "g2 g3 v3"] [--only g2_00,g3_00]
"""
import argparse
import concurrent.futures as cf
import glob
import json
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, 'csrc')
GEN = os.path.join(PKG, '_gen')
OBJ = os.path.join(GEN, 'obj')
LIBDIR = os.path.join(PKG, '_lib')
LIB = os.path.join(LIBDIR, 'libsfx.so')
CLI = os.path.join(LIBDIR, 'sincqed')
KERNELS = os.environ.get('SFX_KERNELS', '/root/reference/kernels')
ARCH = os.environ.get('SFX_ARCH', 'gfx950')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
HIPFLAGS = ['-O3', '--offload-arch=' + ARCH, '-std=c++17', '-fPIC', '-ffp-contract=off',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-I' + CSRC, '-I' + os.path.join(ROOT, 'include')]
DEFAULT_CLASSES = 'g2 g3 v3'


def log(*a):
    print('[build]', *a, flush=True)


def emit_all(classes, only=None):
    sys.path.insert(0, ROOT)
    from sincfeyn_amd.emit import hipgen
    srcs = []
    for c in classes.split():
        srcs += sorted(glob.glob(os.path.join(KERNELS, c, '*.c')))
    if only:
        srcs = [s for s in srcs if os.path.basename(s)[:-2] in only]
    names = []
    for s in srcs:
        out, meta = hipgen.emit(s, GEN)
        names.append(meta['name'])
    return names


def write_registry(names):
    lines = ['// generated by sincfeyn_amd/build.py: the graphs of this build', '#include "sfx_integrand.hpp"']
    lines += ['extern "C" const sfx_graph_entry SFX_E_%s;' % n for n in names]
    lines.append('extern "C" const sfx_graph_entry *const SFX_GRAPHS[] = {')
    lines += ['    &SFX_E_%s,' % n for n in names]
    lines += ['    nullptr };', 'extern "C" const int SFX_NGRAPHS = %d;' % len(names), '']
    path = os.path.join(GEN, 'sfx_registry.hip')
    text = '\n'.join(lines)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, 'w') as f:
            f.write(text)
    return path


TUNE_TABLE = os.path.join(PKG, 'tune_table.json')
# the build variants the tuner (tune.py) measures; a graph's entry in tune_table.json names one
VARIANTS = {
    'base': [],
    'w3': ['-DSFX_LIST_WAVES0=3', '-DSFX_LIST_WAVES1=3'],
    'w4': ['-DSFX_LIST_WAVES0=4', '-DSFX_LIST_WAVES1=3'],
    'burr': ['-mllvm', '-pre-RA-sched=list-burr'],
    'burr_w3': ['-mllvm', '-pre-RA-sched=list-burr', '-DSFX_LIST_WAVES0=3', '-DSFX_LIST_WAVES1=3'],
    'burr_w4': ['-mllvm', '-pre-RA-sched=list-burr', '-DSFX_LIST_WAVES0=4', '-DSFX_LIST_WAVES1=3'],
    # four 64-child slices per wave in the D >= 6 cell-list kernels (points per thread)
    't4': ['-DSFX_LIST_TASKS=4'],
    'w3_t4': ['-DSFX_LIST_WAVES0=3', '-DSFX_LIST_WAVES1=3', '-DSFX_LIST_TASKS=4'],
    'burr_w3_t4': ['-mllvm', '-pre-RA-sched=list-burr', '-DSFX_LIST_WAVES0=3', '-DSFX_LIST_WAVES1=3', '-DSFX_LIST_TASKS=4'],
}


# every integrand unit: the kernels' leading scalar arguments (point and list addresses, the count's address,
# n, st) arrive preloaded in SGPRs, so a wave's first load -- the list count -- does not wait for its own
# address (gfx950 kernarg preloading; the code object keeps the loads for a launcher that does not preload)
KERNEL_FLAGS = ['-mllvm', '-amdgpu-kernarg-preload-count=16']


def unit_flags(src, use_table=True):
    """extra hipcc flags of one translation unit.  The integrands emitted with node_split (their FP64 point
    function as one switch case per forest node) are scheduled bottom-up for register need, which keeps
    one case within the 256 VGPRs of two waves per SIMD (g3_lbl: 244; the default scheduler's 258 leaves
    one wave).  Other graphs take the variant that tune_table.json names for them (measured by tune.py:
    occupancy of the refinement kernels, scheduler).  Scheduling and register allocation only: the same
    instructions on the same operands, so every variant gives the same values bit for bit."""
    meta = src[:-4] + '.json'
    if src.endswith('.hip') and os.path.exists(meta):
        with open(meta) as f:
            if json.load(f).get('node_split'):
                return ['-mllvm', '-pre-RA-sched=list-burr']
    name = os.path.basename(src)
    if use_table and name.startswith('k_') and os.path.exists(TUNE_TABLE):
        with open(TUNE_TABLE) as f:
            v = json.load(f).get('graphs', {}).get(name[2:-4])
        if v:
            return list(VARIANTS[v])
    return []


def compile_one(src, obj, deps, force, flags):
    if not force and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if all(os.path.getmtime(d) <= t for d in [src] + deps):
            return obj, 0.0, ''
    cmd = [HIPCC] + HIPFLAGS + flags + ['-c', src, '-o', obj + '.tmp']
    import time
    t0 = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s' % (src, p.stderr[-4000:]))
    os.replace(obj + '.tmp', obj)
    return obj, time.time() - t0, p.stderr


def build(force=True, jobs=None, classes=DEFAULT_CLASSES, only=None, libdir=None, objdir=None, variant=None):
    """variant (a VARIANTS key): every integrand unit with those flags instead of the tuned table, into
    libdir/objdir (the tuner's trial builds)"""
    libdir = libdir or LIBDIR
    objdir = objdir or OBJ
    lib = os.path.join(libdir, 'libsfx.so')
    os.makedirs(GEN, exist_ok=True)
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(libdir, exist_ok=True)
    if os.path.isdir(KERNELS):
        names = emit_all(classes, only)
        log('emitted %d integrands from %s' % (len(names), KERNELS))
    else:
        names = sorted(os.path.basename(p)[2:-4] for p in glob.glob(os.path.join(GEN, 'k_*.hip')))
        if only:
            names = [n for n in names if n in only]
        missing = [n for n in names if not all(os.path.exists(os.path.join(GEN, 'k_%s.%s' % (n, x))) for x in ('inc', 'json'))]
        if not names or missing:
            raise RuntimeError('cannot build: the sincFeyn kernel sources are not at %s (set SFX_KERNELS to its kernels/ '
                               'directory) and %s holds %s' % (KERNELS, GEN, 'no emitted integrands' if not names else
                               'incomplete emitted integrands (no .inc/.json for %s)' % ', '.join(missing)))
        log('kernel sources absent; compiling the %d emitted integrands already in %s' % (len(names), GEN))
    reg = write_registry(names)
    hdrs = [os.path.join(CSRC, h) for h in ('sfx_integrand.hpp', 'sfx_dd.hpp', 'sfx_qd.hpp', 'sfx_abi_types.h')]
    rt_hdrs = hdrs + [os.path.join(CSRC, h) for h in ('sfx_device.hpp', 'sfx_engine.hpp', 'sfx_grow.inc', 'sfx_scan.inc', 'sfx_ckpt.inc')]
    rt_hdrs.append(os.path.join(ROOT, 'include', 'sfx.h'))
    tdeps = [TUNE_TABLE] if os.path.exists(TUNE_TABLE) and variant is None else []

    def kflags(src):
        if variant is None:
            return KERNEL_FLAGS + unit_flags(src)
        return KERNEL_FLAGS + (unit_flags(src, use_table=False) or list(VARIANTS[variant]))
    units = [(os.path.join(GEN, 'k_%s.hip' % n), os.path.join(objdir, 'k_%s.o' % n), hdrs + [os.path.join(GEN, 'k_%s.inc' % n)] + tdeps)
             for n in names]
    units.append((reg, os.path.join(objdir, 'sfx_registry.o'), hdrs))
    for rt in ('sfx_runtime.hip', 'sfx_hsum.hip'):
        units.append((os.path.join(CSRC, rt), os.path.join(objdir, rt.replace('.hip', '.o')), rt_hdrs))
    # largest sources first, so the long compiles start early
    units.sort(key=lambda u: -os.path.getsize(u[0]))
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(compile_one, s, o, d, force, kflags(s)) for s, o, d in units]
        for f in cf.as_completed(futs):
            o, dt, _ = f.result()
            objs.append(o)
            if dt > 20:
                log('%s: %.0f s' % (os.path.basename(o), dt))
    objs.sort()
    tmp = lib + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('link failed:\n%s' % p.stderr[-4000:])
    os.replace(tmp, lib)
    log('linked %s (%d graphs)' % (os.path.relpath(lib, ROOT), len(names)))
    if variant is not None:
        return lib
    build_cli()
    opcounts(names, jobs, force)
    return lib


def _opcount_one(name, force):
    import json
    inc = os.path.join(GEN, 'k_%s.inc' % name)
    js = os.path.join(OBJ, 'ops_%s.json' % name)
    if not force and os.path.exists(js) and os.path.getmtime(js) >= os.path.getmtime(inc):
        return name, json.load(open(js))
    D = json.load(open(os.path.join(GEN, 'k_%s.json' % name)))['D']
    exe = os.path.join(OBJ, 'opcount_%s' % name)
    cmd = ['g++', '-O1', '-I' + GEN, '-DSFX_INC="k_%s.inc"' % name, '-DSFX_F=f_%s' % name, '-DSFX_D=%d' % D,
           os.path.join(PKG, 'emit', 'opcount.cpp'), '-o', exe]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('opcount build failed for %s:\n%s' % (name, p.stderr[-2000:]))
    p = subprocess.run([exe], capture_output=True, text=True, check=True)
    d = json.loads(p.stdout)
    with open(js, 'w') as f:
        json.dump(d, f)
    os.remove(exe)
    return name, d


def opcounts(names, jobs, force):
    """per-point FP64 operation counts of every integrand (emit/opcount.cpp) -> _gen/opcounts.json"""
    import json
    out = {}
    with cf.ThreadPoolExecutor(jobs) as ex:
        for name, d in ex.map(lambda n: _opcount_one(n, force), names):
            out[name] = d
    with open(os.path.join(GEN, 'opcounts.json'), 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
    log('per-point op counts of %d integrands -> _gen/opcounts.json' % len(out))


def build_cli():
    src = os.path.join(ROOT, 'src', 'sincqed.c')
    if not os.path.exists(src):
        return None
    cmd = ['gcc', '-O2', '-std=c11', '-Wall', '-I' + os.path.join(ROOT, 'include'), '-o', CLI + '.tmp', src,
           '-L' + LIBDIR, '-lsfx', '-Wl,-rpath,$ORIGIN', '-lm']
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('sincqed CLI build failed:\n%s' % p.stderr[-4000:])
    os.replace(CLI + '.tmp', CLI)
    log('built %s' % os.path.relpath(CLI, ROOT))
    # sincFeyn's GPU back-end ABI (driver/gpu.h) over the library: what its host driver links instead of
    # driver/gpu.cu
    gsrc, glib = os.path.join(ROOT, 'src', 'sfx_gpu.c'), os.path.join(LIBDIR, 'libsfxgpu.so')
    if os.path.exists(gsrc):
        cmd = ['gcc', '-O2', '-std=c11', '-Wall', '-fPIC', '-shared', '-I' + os.path.join(ROOT, 'include'), '-o', glib + '.tmp',
               gsrc, '-L' + LIBDIR, '-lsfx', '-Wl,-rpath,$ORIGIN']
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError('libsfxgpu build failed:\n%s' % p.stderr[-4000:])
        os.replace(glib + '.tmp', glib)
        log('built %s' % os.path.relpath(glib, ROOT))
    return CLI


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--force', action='store_true')
    ap.add_argument('--jobs', type=int, default=None)
    ap.add_argument('--classes', default=DEFAULT_CLASSES)
    ap.add_argument('--only', default=None, help='comma-separated graph names')
    a = ap.parse_args(argv)
    only = set(a.only.split(',')) if a.only else None
    build(force=a.force, jobs=a.jobs, classes=a.classes, only=only)


if __name__ == '__main__':
    main()
