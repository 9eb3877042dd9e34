// Host-only: the optimal alphabetic prefix code that keys round 0 of large sorts (dcx_engine.hip).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace dcx {

// Optimal alphabetic prefix code of the weights w[0, m): the binary tree with the symbols as leaves in
// order that minimizes the weighted depth (Hu-Tucker's problem), by the dynamic program over contiguous
// symbol ranges with the monotone split bounds of Knuth and Yao (root(i, j - 1) <= root(i, j) <=
// root(i + 1, j)), O(m^2).  Any split gives a valid alphabetic code; the bounds only make it fast.
// len[s] and code[s] (the codeword's len[s] bits, MSB first) per symbol; returns the longest length, or 0
// when a codeword would exceed max_len bits.
inline uint32_t alphabetic_code(const std::vector<double>& w, uint32_t max_len, std::vector<uint32_t>& len,
                                std::vector<uint32_t>& code) {
    const int m = (int)w.size();
    len.assign(m, 0);
    code.assign(m, 0);
    if (m == 0) return 0;
    if (m == 1) {
        len[0] = 1;
        return 1;
    }
    std::vector<double> pre(m + 1, 0.0);
    for (int i = 0; i < m; i++) pre[i + 1] = pre[i] + w[i];
    std::vector<double> cost((size_t)m * m, 0.0);
    std::vector<int> root((size_t)m * m, 0);
    auto at = [m](int i, int j) { return (size_t)i * m + j; };
    for (int i = 0; i < m; i++) root[at(i, i)] = i;
    for (int d = 1; d < m; d++)
        for (int i = 0; i + d < m; i++) {
            const int j = i + d;
            // split after k: leaves [i, k] left, [k + 1, j] right
            const int lo = std::max(i, root[at(i, j - 1)]);
            const int hi = std::max(lo, std::min(j - 1, root[at(i + 1, j)]));
            double best = 0;
            int bk = -1;
            for (int k = lo; k <= hi; k++) {
                const double c = cost[at(i, k)] + cost[at(k + 1, j)];
                if (bk < 0 || c < best) {
                    best = c;
                    bk = k;
                }
            }
            cost[at(i, j)] = best + (pre[j + 1] - pre[i]);
            root[at(i, j)] = bk;
        }
    struct Node {
        int i, j;
        uint32_t depth, prefix;
    };
    std::vector<Node> st{{0, m - 1, 0, 0}};
    uint32_t longest = 0;
    while (!st.empty()) {
        const Node nd = st.back();
        st.pop_back();
        if (nd.i == nd.j) {
            len[nd.i] = nd.depth;
            code[nd.i] = nd.prefix;
            longest = std::max(longest, nd.depth);
            continue;
        }
        if (nd.depth >= max_len) return 0;
        const int k = root[at(nd.i, nd.j)];
        st.push_back({nd.i, k, nd.depth + 1, nd.prefix << 1});
        st.push_back({k + 1, nd.j, nd.depth + 1, (nd.prefix << 1) | 1u});
    }
    return longest;
}

}  // namespace dcx
