"""The reference-side binding in INTEGRATION.md, checked on the CPU.

1. Its Rust `extern "C"` block declares exactly the symbols of
   include/soil_hip.h, each with the header's number of arguments.
2. Its plugins register every HIP system at the phase, label, `before` and
   `run_if` of the reference system it replaces. The table below transcribes
   the reference registrations (file:line cited per row); where the reference
   tree is present its plugin sources are parsed and must agree with the table.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

# hip system -> (reference system, phase, label, before, run_if); `before`
# names the HIP system standing in for the reference one
REGISTRATIONS = {
    # crates/soil_core/src/neighbor.rs:549-555
    "hip_decide_rebuild": ("decide_rebuild", "PostInitialIntegration", "decide_rebuild", "hip_remove_ghost_atoms",
                           "in_state(CommState::CommunicateOnly)"),
    # crates/soil_core/src/neighbor.rs:556-561
    "hip_sort_atoms_by_bin": ("sort_atoms_by_bin", "PreNeighbor", "sort_atoms", "hip_borders", None),
    # crates/soil_core/src/neighbor.rs:562-565
    "hip_bin_neighbor_list": ("bin_neighbor_list", "Neighbor", None, None, "in_state(CommState::FullRebuild)"),
    # crates/soil_core/src/comm.rs:185-188
    "hip_borders": ("borders", "PreNeighbor", None, None, "in_state(CommState::FullRebuild)"),
    # crates/soil_core/src/comm.rs:189-192
    "hip_forward_comm_borders": ("forward_comm_borders", "PreNeighbor", None, None,
                                 "in_state(CommState::CommunicateOnly)"),
    # crates/soil_core/src/comm.rs:193-196
    "hip_reverse_send_force": ("reverse_send_force", "PostForce", "reverse_send_force", None, "newton_on"),
    # crates/soil_core/src/atom.rs:615-618
    "hip_remove_ghost_atoms": ("remove_ghost_atoms", "PostInitialIntegration", None, None,
                               "in_state(CommState::FullRebuild)"),
    # crates/soil_core/src/atom.rs:619
    "hip_zero_all_forces": ("zero_all_forces", "PostInitialIntegration", None, None, None),
    # crates/soil_core/src/domain.rs:442-445
    "hip_shrink_wrap": ("shrink_wrap", "PreExchange", "shrink_wrap", '"pbc"', None),
    # crates/soil_core/src/domain.rs:446-449
    "hip_pbc": ("pbc", "PreExchange", "pbc", None, "in_state(CommState::FullRebuild)"),
    # crates/soil_verlet/src/lib.rs:113-114
    "hip_initial_integration": ("initial_integration", "InitialIntegration", None, None, None),
    "hip_final_integration": ("final_integration", "FinalIntegration", None, None, None),
    # crates/soil_deform/src/lib.rs:268
    "hip_apply_deform": ("apply_deform", "PreInitialIntegration", None, None, None),
}
REF_FILES = ["crates/soil_core/src/neighbor.rs", "crates/soil_core/src/comm.rs", "crates/soil_core/src/atom.rs",
             "crates/soil_core/src/domain.rs", "crates/soil_verlet/src/lib.rs", "crates/soil_deform/src/lib.rs"]


def _strip_c_comments(t):
    return re.sub(r"/\*.*?\*/", "", t, flags=re.S)


def header_arity():
    t = _strip_c_comments(open(os.path.join(ROOT, "include", "soil_hip.h")).read())
    out = {}
    for name, args in re.findall(r"\b(soil_hip_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", t):
        a = args.strip()
        out[name] = 0 if a in ("", "void") else len(a.split(","))
    return out


def doc_rust_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```rust\n(.*?)```", text, flags=re.S)


def doc_extern_arity():
    blocks = [b for b in doc_rust_blocks() if 'extern "C" {' in b]
    assert len(blocks) == 1
    body = blocks[0].split('extern "C" {', 1)[1]
    out = {}
    for name, args in re.findall(r"pub fn (soil_hip_[a-z0-9_]+)\((.*?)\)\s*->", body, flags=re.S):
        # split on top-level commas only (fn-pointer types hold their own)
        depth, n, cur = 0, 0, ""
        for ch in args:
            depth += ch in "(<[" and 1 or 0
            depth -= ch in ")>]" and 1 or 0
            if ch == "," and depth == 0:
                n += 1 if cur.strip() else 0
                cur = ""
            else:
                cur += ch
        n += 1 if cur.strip() else 0
        assert name not in out, f"{name} declared twice"
        out[name] = n
    return out


def parse_registrations(code):
    """system -> (phase, label, before, run_if) for every add_update_system(...) call."""
    regs = {}
    for m in re.finditer(r"add_update_system\(\s*([a-z_][a-z_0-9:]*)(.*?),\s*ParticleSimScheduleSet::(\w+)\s*,?\s*\)",
                         code, flags=re.S):
        sysname = m.group(1).split("::")[-1]
        mods = m.group(2)
        label = re.search(r'\.label\("([^"]+)"\)', mods)
        before = re.search(r"\.before\(([^)]*)\)", mods)
        run_if = re.search(r"\.run_if\((.*)\)\s*$", mods.strip(), flags=re.S)
        b = before.group(1).strip() if before else None
        if b and "::" in b:
            b = b.split("::")[-1]
        regs.setdefault(sysname, []).append((m.group(3), label.group(1) if label else None, b,
                                             run_if.group(1).strip() if run_if else None))
    return regs


def test_extern_block_declares_every_header_symbol_with_its_arity():
    h = header_arity()
    d = doc_extern_arity()
    assert len(h) >= 70
    assert set(d) == set(h), (sorted(set(h) - set(d)), sorted(set(d) - set(h)))
    for name in h:
        assert d[name] == h[name], f"{name}: {d[name]} arguments in INTEGRATION.md, {h[name]} in the header"


def test_extern_block_declares_the_types_it_uses():
    code = "\n".join(doc_rust_blocks())
    for ty in ("c_void", "c_char", "c_int"):
        assert re.search(r"use std::os::raw::\{[^}]*\b%s\b" % ty, code), ty
    for st in ("soil_hip_ctx", "soil_hip_box", "soil_hip_batch", "soil_hip_server", "soil_hip_box_desc",
               "soil_hip_device_view", "soil_hip_restart"):
        assert re.search(r"pub struct %s\b" % st, code), st
    assert "pub type soil_hip_force_hook" in code


def test_plugin_registrations_match_the_reference_table():
    regs = parse_registrations("\n".join(doc_rust_blocks()))
    for hip, (_, phase, label, before, run_if) in REGISTRATIONS.items():
        assert hip in regs, f"{hip} is not registered"
        assert regs[hip] == [(phase, label, before, run_if)], (hip, regs[hip])
    # no HIP system is registered at a phase the table does not give it
    assert set(k for k in regs if k.startswith("hip_")) == set(REGISTRATIONS)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_reference_table_matches_the_reference_sources():
    regs = {}
    for f in REF_FILES:
        for k, v in parse_registrations(open(os.path.join(REF, f)).read()).items():
            regs.setdefault(k, []).extend(v)
    for hip, (ref, phase, label, before, run_if) in REGISTRATIONS.items():
        assert ref in regs, f"reference registers no {ref}"
        # the plugin's registration is among the file's (its test modules register some systems too)
        want_before = before.replace("hip_", "") if before and before.startswith("hip_") else before
        assert (phase, label, want_before, run_if) in regs[ref], (ref, regs[ref])
