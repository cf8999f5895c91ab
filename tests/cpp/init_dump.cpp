// Writes the initial state that the command line (include/cellato/experiments/initial_state.hpp)
// generates, as raw int32 enum values, to stdout: `init_dump <automaton> <width> <height> <seed>`.
// tests/test_cli.py compares it with the oracle's restatement of the reference generators.
#include <cstdio>
#include <cstdlib>

#include "cellato/experiments/initial_state.hpp"

int main(int argc, char** argv) {
    if (argc != 5) return 2;
    const auto v = cellato::experiments::initial_state(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    return std::fwrite(v.data(), sizeof(v[0]), v.size(), stdout) == v.size() ? 0 : 1;
}
