"""Drop-in check: the reference's own rule headers (src/<automaton>/algorithm.hpp) compile
against this repository's rule vocabulary (include/cellato/core/ast.hpp) and produce the same
rule types as include/cellato/automata/*.hpp, node for node; the dictionaries list the states
in the reference's order (src/<automaton>/config.hpp). Needs the reference checkout, so it is
skipped where that is absent."""
import re
import subprocess
from pathlib import Path

import pytest

REF = Path("/root/reference")
REPO = Path(__file__).resolve().parent.parent
AUTOMATA = {  # directory -> (namespace, algorithm alias)
    "game_of_life": ("game_of_life", "gol_algorithm"), "fire": ("fire", "fire_algorithm"),
    "wire": ("wire", "wire_algorithm"), "excitable": ("excitable", "ghm_algorithm"),
    "brian": ("brian", "brian_algorithm"), "maze": ("maze", "maze_algorithm"),
    "traffic": ("traffic", "traffic_algorithm"), "fluid": ("fluid", "fluid_algorithm"),
    "critters": ("critters", "critters_algorithm"), "cyclic": ("cyclic", "cyclic_algorithm"),
}

PRELUDE = r"""
#include <type_traits>
#include "cellato/automata/all.hpp"
using namespace cellato::ast;
// structural equality of rule types; state literals compare by value
template <class A, class B> struct same_rule : std::is_same<A, B> {};
template <auto a, auto b> struct same_rule<state_constant<a>, state_constant<b>>
    : std::bool_constant<(long long)a == (long long)b> {};
template <class A, class B, int k> struct same_rule<has_bit_set<A, k>, has_bit_set<B, k>> : same_rule<A, B> {};
template <template <class...> class N, class... As, class... Bs> struct same_rule<N<As...>, N<Bs...>>
    : std::bool_constant<sizeof...(As) == sizeof...(Bs) && (same_rule<As, Bs>::value && ...)> {};
"""


@pytest.mark.skipif(not REF.exists(), reason="reference checkout not present")
def test_reference_rules_compile_unchanged(tmp_path):
    src = [PRELUDE, "namespace ref {"]
    for d in AUTOMATA:
        src.append(f'#include "{REF}/src/{d}/algorithm.hpp"')
    src.append("}")
    for d, (ns, alg) in AUTOMATA.items():
        src.append(f"static_assert(same_rule<ref::{ns}::{alg}, ::{ns}::{alg}>::value, \"{d}\");")
        cfg = (REF / "src" / d / "config.hpp").read_text()
        m = re.search(r"(int_based_state_dictionary<\s*(\w+)\s*>|state_dictionary<(.*?)>;)", cfg, re.S)
        if m.group(2):
            src.append(f"static_assert(std::is_same_v<::{ns}::config::state_dictionary, "
                       f"cellato::memory::grids::int_based_state_dictionary<{m.group(2) if m.group(2).isdigit() else '::' + ns + '::' + m.group(2)}>>);")
        else:
            names = [n.split("::")[-1].strip() for n in m.group(3).split(",")]
            for i, n in enumerate(names):
                src.append(f"static_assert(::{ns}::config::state_dictionary::state_to_index("
                           f"::{ns}::config::cell_state::{n}) == {i}, \"{d} dictionary order\");")
    f = tmp_path / "dropin.cpp"
    f.write_text("\n".join(src) + "\nint main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", f"-I{REPO}/include", f"-I{REPO}/include/cellato", str(f)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
