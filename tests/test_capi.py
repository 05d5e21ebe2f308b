"""The C-ABI library loads on a CPU-only host and exports every entry point
that include/*.h declares (no compute calls: there may be no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "hptt_amd", "libhptt_amd.so")


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(hpttx_\w+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import hptt_amd.build
        hptt_amd.build.build()
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    syms = declared_symbols()
    assert len(syms) >= 15, syms
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(\w+)$", out, flags=re.M))
    missing = sorted(s for s in syms if s not in exported)
    assert not missing, missing
    for s in syms:
        getattr(lib, s)


def test_python_binding_loads_and_validates(lib):
    import hptt_amd
    L = hptt_amd.lib()
    assert L.hpttx_version() == 100
    # Validation happens before any device work, so it runs without a GPU:
    # each reference nullptr case (src/hptt.cc:25-56) is refused with a status.
    p = ctypes.c_void_p()
    u = hptt_amd._u32
    fake = 0x100000
    cases = [
        (0, fake, [4, 4], [1, 0], None, None, hptt_amd.EINVAL),     # null A
        (fake, fake, [4], [0], None, None, hptt_amd.EINVAL),        # rank 1
        (fake, fake, [4, 4], [1, 1], None, None, hptt_amd.EINVAL),  # bad perm
        (fake, fake, [4, 0], [1, 0], None, None, hptt_amd.EINVAL),  # zero size
        (fake, fake, [4, 4], [1, 0], [3, 4], None, hptt_amd.EINVAL),
        (fake, fake, [4, 4], [1, 0], None, [4, 3], hptt_amd.EINVAL),
        (fake, fake, [2] * 7, list(range(7)), None, None, hptt_amd.EUNSUPPORTED),
    ]
    for a, b, size, perm, io, oo, want in cases:
        rc = L.hpttx_plan_create_s(a or None, b, len(perm), u(size), u(perm), 1.0, 0.0, 0,
                                   0.0, u(io), u(oo), None, ctypes.byref(p))
        assert rc == want, (size, perm, rc, hptt_amd.last_error())
        assert not p.value
        assert hptt_amd.last_error()
    # the drop-in returns None (nullptr) for the same cases
    assert hptt_amd.create_plan(fake, fake, [4], [0], 1.0, 0.0, dtype="s") is None


def test_python_structs_match_the_c_abi(tmp_path):
    """Every ctypes Structure of the Python binding has the size and field
    offsets of its struct in include/hptt_amd.h (measured by the host C
    compiler), so the binding cannot read or write past a field."""
    import hptt_amd
    pairs = [(hptt_amd.PlanInfo, "hpttx_plan_info"), (hptt_amd.BatchItem, "hpttx_batch_item"),
             (hptt_amd.BatchItemInfo, "hpttx_batch_item_info"),
             (hptt_amd.BatchInfo, "hpttx_batch_info"),
             (hptt_amd.ServerConfigC, "hpttx_server_config"),
             (hptt_amd.Completion, "hpttx_completion"), (hptt_amd.ServerStats, "hpttx_server_stats")]
    cname = lambda f: "in" if f == "in_" else f
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hptt_amd.h"', 'int main(void) {']
    for cls, cs in pairs:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cs, cs))
        for f, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cs, f, cs, cname(f)))
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cls, cs in pairs:
        assert got[(cs, "size")] == ctypes.sizeof(cls), cs
        for f, _ in cls._fields_:
            assert got[(cs, f)] == getattr(cls, f).offset, (cs, f)
    # and the same fields, in order (a missing trailing field could hide in padding)
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "hptt_amd.h")).read(),
                  flags=re.S)
    for cls, cs in pairs:
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cs, cs), text, flags=re.S).group(1)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            head, _, rest = decl.partition(" ")
            while rest.startswith(("int", "uint", "double", "void", "*", "const")) and " " in rest:
                head, _, rest = rest.partition(" ")
            names += [re.sub(r"[\s*]|\[.*\]", "", n) for n in rest.split(",")]
        assert [cname(f) for f, _ in cls._fields_] == names, (cs, names)


def test_cpp_dropin_header_compiles(tmp_path):
    """include/hptt/hptt.h compiles as plain C++17 against the C ABI header
    (host compiler, no HIP headers needed)."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "hptt/hptt.h"\n'
                   'int main() { auto p = hptt::create_plan<float>(nullptr, nullptr, {2, 2},'
                   ' {1, 0}, 1.f, 0.f, 1); return p ? 1 : 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


# Where the reference's public headers are, on a machine that has its sources
# (elsewhere the test skips).
REFERENCE_INC = "/root/reference/inc"

# The mangled factory names the reference's front end looks up in its plugin
# (src/hptt.cc:71-84).
REFERENCE_FACTORIES = [
    "_Z24create_trans_plan_impl_sPKfPfjRKSt6vectorIjSaIjEES6_ffjdS6_S6_",
    "_Z24create_trans_plan_impl_dPKdPdjRKSt6vectorIjSaIjEES6_ddjdS6_S6_",
    "_Z24create_trans_plan_impl_cPKCfPS_jRKSt6vectorIjSaIjEES7_ffjdS7_S7_",
    "_Z24create_trans_plan_impl_zPKCdPS_jRKSt6vectorIjSaIjEES7_ddjdS7_S7_",
]


@pytest.mark.skipif(not os.path.isdir(REFERENCE_INC), reason="reference headers not present")
def test_reference_plugin_shim_exports_the_factories(tmp_path):
    """tests/cpp/plugin_shim.cpp, built against the reference's own public
    headers and linked to the library, exports exactly the four factories the
    reference's front end dlsyms, so the unchanged front end can load it."""
    so = tmp_path / "libhptt_avx2.so"
    r = subprocess.run(["g++", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I", REFERENCE_INC,
                        "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "plugin_shim.cpp"),
                        "-L", os.path.join(ROOT, "hptt_amd"), "-lhptt_amd", "-o", str(so)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True)
    assert nm.returncode == 0, nm.stderr
    exported = {line.split()[-1] for line in nm.stdout.splitlines() if line.strip()}
    for name in REFERENCE_FACTORIES:
        assert name in exported, name


def test_python_shape_checks_before_the_c_abi():
    """The binding refuses arrays of the wrong length before the C ABI reads
    rank entries of each (the reference returns nullptr, src/hptt.cc:35-39),
    and computes the spans tensors must hold."""
    import hptt_amd
    fake = 0x100000
    bad = [
        dict(in_size=[4, 4, 4], perm=[1, 0]),
        dict(in_size=[4], perm=[1, 0]),
        dict(in_size=[4, 4], perm=[1, 0], in_outer_size=[5]),
        dict(in_size=[4, 4], perm=[1, 0], out_outer_size=[5, 5, 5]),
    ]
    for kw in bad:
        assert hptt_amd.create_plan(fake, fake, alpha=1.0, beta=0.0, dtype="s", **kw) is None
        with pytest.raises(hptt_amd.HpttError) as ei:
            hptt_amd.create_plan_checked(fake, fake, alpha=1.0, beta=0.0, dtype="s", **kw)
        assert ei.value.status == hptt_amd.EINVAL
        with pytest.raises(hptt_amd.HpttError):
            hptt_amd.make_batch_item(fake, fake, kw["in_size"], kw["perm"], 1.0, 0.0, dtype="s",
                                     in_outer_size=kw.get("in_outer_size", ()),
                                     out_outer_size=kw.get("out_outer_size", ()))
    with pytest.raises(hptt_amd.HpttError) as ei:
        hptt_amd.make_batch_item(fake, fake, [2] * 7, list(range(7)), 1.0, 0.0, dtype="s")
    assert ei.value.status == hptt_amd.EUNSUPPORTED
    # spans: column-major over the outer sizes (inc/hptt/tensor.tcc:186-196)
    assert hptt_amd._spans([3, 4], [1, 0], (), ()) == [12, 12]
    assert hptt_amd._spans([3, 4], [1, 0], [5, 4], [6, 3]) == [1 + 2 + 3 * 5, 1 + 3 + 2 * 6]
    assert hptt_amd._spans([3, 4], [0, 0], (), ()) is None


def test_server_completion_bookkeeping_host():
    """tests/cpp/serve_book_test: a failed slot's requests are marked failed
    with the reason and counted, a completed slot's done (no GPU used)."""
    exe = os.path.join(ROOT, "tests", "cpp", "serve_book_test")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"),
                        "serve_book_test"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "SERVE BOOK OK" in r.stdout, r.stdout + r.stderr


def test_noalloc_entry_point_source_calls_no_allocator():
    """hpttx_batch_run_noalloc's source (and the batch launch helpers it
    calls) names no device or pinned-host allocation routine."""
    src = open(os.path.join(ROOT, "hptt_amd", "csrc", "noalloc.hip")).read()
    code = re.sub(r"//[^\n]*", "", src)
    for name in ("hipMalloc", "hipFree", "hipHostMalloc", "hipMallocAsync", "hipExtMalloc",
                 "hipMallocManaged", "hipEventCreate", "hipStreamCreate"):
        assert name not in code, name
    assert "static __device__ __align__(256) unsigned char scratch[" in code
    assert "hipGetSymbolAddress" in code
