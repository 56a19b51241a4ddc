// Iteration order of std::unordered_map<int, ...> after reserve(n) and inserts of the given
// keys in order (test infrastructure: the oracle's restatement of the reference's 1+1D
// adjacency, fespace/adjacency/adjacency_1d.hpp:66-124, groups faces in such a map, and the
// order in which it visits the groups decides the order of the grading closure's
// refinements).  Built with the host compiler by oracle/adaptive1d.py.
#include <cstddef>
#include <cstdint>
#include <unordered_map>
#include <vector>

extern "C" int64_t oracle_hash_order(int64_t reserve_n, int64_t n_keys, const int* keys, int* out)
{
    std::unordered_map<int, std::vector<int>> m;
    m.reserve(static_cast<size_t>(reserve_n));
    for (int64_t i = 0; i < n_keys; ++i) m[keys[i]].push_back(static_cast<int>(i));
    int64_t k = 0;
    for (const auto& kv : m) out[k++] = kv.first;
    return k;
}
