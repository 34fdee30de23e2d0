// Locality order of the nodes for the solver's sparse kernels.
//
// The block SpMM gathers, for every row, the 64-column rows of X of its
// neighbours (512 bytes each in complex FP32).  How often such a row is
// fetched from beyond L2 depends on how far apart in row order its
// neighbours are, so the solver works on a relabelled copy of the system:
// nodes sorted along a Morton (Z-order) curve of their coordinates, which
// keeps spatially close nodes close in row order whatever order the mesh
// file or generator used.  Only the solver's private copy is relabelled; the
// mesh, its CSR pattern and the assembled matrix keep the reference order.
#include <algorithm>
#include <array>
#include <iterator>
#include <cstdint>
#include <numeric>
#include <vector>

#include "krylov.hpp"
#include "mesh.hpp"

namespace tmi {

// spread the low 21 bits of v to every third bit (3D) / 32 bits to every
// second bit (2D)
static uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
static uint64_t spread2(uint64_t v)
{
    v &= 0xffffffffULL;
    v = (v | v << 16) & 0x0000ffff0000ffffULL;
    v = (v | v << 8) & 0x00ff00ff00ff00ffULL;
    v = (v | v << 4) & 0x0f0f0f0f0f0f0f0fULL;
    v = (v | v << 2) & 0x3333333333333333ULL;
    v = (v | v << 1) & 0x5555555555555555ULL;
    return v;
}

std::vector<int> morton_order(int dim, int n, const double *x)
{
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
        lo[d] = hi[d] = x[d];
        for (int i = 1; i < n; i++) {
            lo[d] = std::min(lo[d], x[(size_t)i * dim + d]);
            hi[d] = std::max(hi[d], x[(size_t)i * dim + d]);
        }
    }
    // one scale for all axes keeps the curve's cells cubic
    double ext = 0.0;
    for (int d = 0; d < dim; d++) ext = std::max(ext, hi[d] - lo[d]);
    const double levels = dim == 3 ? 2097151.0 : 4294967295.0;
    const double s = ext > 0.0 ? levels / ext : 0.0;
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; i++) {
        uint64_t k = 0;
        for (int d = 0; d < dim; d++) {
            double q = (x[(size_t)i * dim + d] - lo[d]) * s;
            uint64_t c = (uint64_t)std::min(levels, std::max(0.0, q));
            k |= (dim == 3 ? spread3(c) : spread2(c)) << d;
        }
        key[i] = k;
    }
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    return perm;
}


// Rows in groups of `group` (a power of two) that share most of their
// columns, for the row-block SpMM: a block of consecutive rows gathers each
// distinct neighbour row once, so the fewer distinct columns a block's rows
// have together, the fewer gathers.  Heavy-edge matching, log2(group)
// rounds: the groups (single rows at first, visited in the order `seed`)
// are paired, each unpaired group with the unpaired neighbouring group whose
// column union shares most columns with its own (ties: the group visited
// first); a pair's union is the union of the two.  The order lists the
// groups that reached full size in visiting order, then the others.  On the
// 250k-node tetrahedral ellipsoid, blocks of 8 rows then hold 0.41 distinct
// columns per nonzero against 0.54 for 8 consecutive Morton rows.
std::vector<int> cluster_order(int n, const std::vector<int> &rp, const std::vector<int> &ci,
                               const std::vector<int> &seed, int group)
{
    std::vector<std::vector<int>> grp(n), uni(n);
    std::vector<int> owner(n);
    for (int g = 0; g < n; g++) {
        const int v = seed[g];
        grp[g].assign(1, v);
        uni[g].assign(ci.begin() + rp[v], ci.begin() + rp[v + 1]);
        std::sort(uni[g].begin(), uni[g].end());
        owner[v] = g;
    }
    auto shared = [](const std::vector<int> &a, const std::vector<int> &b) {
        int c = 0;
        for (size_t i = 0, j = 0; i < a.size() && j < b.size();) {
            if (a[i] < b[j]) i++;
            else if (b[j] < a[i]) j++;
            else { c++; i++; j++; }
        }
        return c;
    };
    std::vector<int> seen;
    for (int size = 1; size < group; size *= 2) {
        const int ng = (int)grp.size();
        std::vector<int> match(ng, -1);
        seen.assign(ng, -1);
        std::vector<std::vector<int>> ngrp, nuni;
        ngrp.reserve(ng / 2 + 1);
        nuni.reserve(ng / 2 + 1);
        for (int g = 0; g < ng; g++) {
            if (match[g] >= 0) continue;
            int best = -1, bw = -1;
            for (int v : uni[g]) {
                const int h = owner[v];
                if (h == g || match[h] >= 0 || seen[h] == g) continue;
                seen[h] = g;
                const int w = shared(uni[g], uni[h]);
                if (w > bw || (w == bw && h < best)) { best = h; bw = w; }
            }
            match[g] = (int)ngrp.size();
            if (best >= 0) {
                match[best] = (int)ngrp.size();
                std::vector<int> m(grp[g]);
                m.insert(m.end(), grp[best].begin(), grp[best].end());
                std::vector<int> u;
                u.reserve(uni[g].size() + uni[best].size());
                std::set_union(uni[g].begin(), uni[g].end(), uni[best].begin(), uni[best].end(), std::back_inserter(u));
                ngrp.push_back(std::move(m));
                nuni.push_back(std::move(u));
            } else {
                ngrp.push_back(std::move(grp[g]));
                nuni.push_back(std::move(uni[g]));
            }
        }
        grp.swap(ngrp);
        uni.swap(nuni);
        for (int g = 0; g < (int)grp.size(); g++)
            for (int v : grp[g]) owner[v] = g;
    }
    std::vector<int> perm;
    perm.reserve(n);
    for (int pass = 0; pass < 2; pass++)
        for (const auto &g : grp)
            if (((int)g.size() == group) == (pass == 0)) perm.insert(perm.end(), g.begin(), g.end());
    return perm;
}

// Greedy partition of the rows (in order) into blocks whose distinct
// columns plus rows number at most umax and whose nonzeros number at most
// nzmax: brow / bu bound the blocks' rows and their entries of ucols;
// lcol[k] is the position of colidx[k] in its block's list.  Returns false
// if a single row does not fit.
static bool row_blocks(int n, const std::vector<int> &rp, const std::vector<int> &ci, int umax, int nzmax,
                       std::vector<int> &brow,
                       std::vector<int> &bu, std::vector<int> &ucols, std::vector<unsigned short> &lcol)
{
    std::vector<int> mark(n, -1), pos(n, 0);
    brow.assign(1, 0);
    bu.assign(1, 0);
    ucols.clear();
    lcol.assign(ci.size(), 0);
    int blk = 0, cur = 0;
    for (int r = 0; r < n; r++) {
        if (rp[r + 1] - rp[r] + 1 > umax || rp[r + 1] - rp[r] > nzmax) return false;
        int fresh = 0;
        for (int k = rp[r]; k < rp[r + 1]; k++) fresh += mark[ci[k]] != blk;
        const int rows = r - brow.back();
        if (cur + fresh + rows + 1 > umax || rp[r + 1] - rp[brow.back()] > nzmax) {
            brow.push_back(r);
            bu.push_back((int)ucols.size());
            blk++;
            cur = 0;
        }
        for (int k = rp[r]; k < rp[r + 1]; k++) {
            const int c = ci[k];
            if (mark[c] != blk) {
                mark[c] = blk;
                pos[c] = cur++;
                ucols.push_back(c);
            }
            lcol[k] = (unsigned short)pos[c];
        }
    }
    brow.push_back(n);
    bu.push_back((int)ucols.size());
    return true;
}

// Row blocks of RB_ROWS consecutive rows: the ascending union of their
// columns with a row mask per column, and the CSR slots of the block's
// coefficients in (column, row) order.  The block's coefficients occupy the
// block's own CSR range, so the kernel finds them at rowptr[first row].
static void row_block_unions(int n, const std::vector<int> &rp, const std::vector<int> &ci, std::vector<int> &ofs,
                             std::vector<int> &col, std::vector<rb_mask_t> &mask, std::vector<int> &slot)
{
    const int nb = (n + RB_ROWS - 1) / RB_ROWS;
    ofs.assign(1, 0);
    col.clear();
    mask.clear();
    slot.assign(ci.size(), 0);
    std::vector<std::array<int, 3>> e;   // (column, row in block, CSR slot)
    for (int b = 0; b < nb; b++) {
        const int r0 = b * RB_ROWS, r1 = std::min(n, r0 + RB_ROWS);
        e.clear();
        for (int r = r0; r < r1; r++)
            for (int k = rp[r]; k < rp[r + 1]; k++) e.push_back({ci[k], r - r0, k});
        std::sort(e.begin(), e.end());
        int at = rp[r0];
        for (size_t i = 0; i < e.size(); i++) {
            if (i == 0 || e[i][0] != e[i - 1][0]) {
                col.push_back(e[i][0]);
                mask.push_back(0);
            }
            mask.back() |= (rb_mask_t)(1u << e[i][1]);
            slot[at++] = e[i][2];
        }
        ofs.push_back((int)col.size());
    }
}

const Mesh::NodeOrder &Mesh::build_order(int kind, hipStream_t st)
{
    TMI_CHECK(kind == ORDER_NATURAL || kind == ORDER_MORTON || kind == ORDER_CLUSTER, TMI_ERR_INVALID_ARG,
              "unknown node order %d", kind);
    build_csr(st);
    NodeOrder &order = this->order[kind];
    if (order.valid) return order;
    std::vector<int> perm;
    const std::vector<int> rp = rowptr.to_host(st), ci = colidx.to_host(st);
    if (kind == ORDER_MORTON || kind == ORDER_CLUSTER) perm = morton_order(dim, n, h_nodes.data());
    else {
        perm.resize(n);
        std::iota(perm.begin(), perm.end(), 0);
    }
    if (kind == ORDER_CLUSTER) perm = cluster_order(n, rp, ci, perm, RB_ROWS);
    std::vector<int> iperm(n);
    for (int r = 0; r < n; r++) iperm[perm[r]] = r;
    std::vector<int> prp(n + 1), pci(nnz), slot(nnz), pdiag(n, -1);
    std::vector<std::pair<int, int>> row;
    prp[0] = 0;
    for (int r = 0; r < n; r++) {
        const int o = perm[r];
        row.clear();
        for (int k = rp[o]; k < rp[o + 1]; k++) row.emplace_back(iperm[ci[k]], k);
        std::sort(row.begin(), row.end());
        int at = prp[r];
        for (auto &e : row) {
            if (e.first == r) pdiag[r] = at;
            pci[at] = e.first;
            slot[at] = e.second;
            at++;
        }
        prp[r + 1] = at;
    }
    order.perm.upload(perm.data(), n, st);
    // 4 trailing copies of nnz: the LDS-DMA SpMM stages row pointers in 16-byte chunks
    prp.insert(prp.end(), 4, prp[n]);
    order.rowptr.upload(prp.data(), n + 5, st);
    order.colidx.upload(pci.data(), nnz, st);
    order.slot.upload(slot.data(), nnz, st);
    order.diagpos.upload(pdiag.data(), n, st);
    std::vector<int> brow, bu, ucols;
    std::vector<unsigned short> lcol;
    order.nb = 0;
    if (row_blocks(n, prp, pci, LDS_ROWS_C32, LDS_NZ_C32, brow, bu, ucols, lcol)) {
        order.nb = (int)brow.size() - 1;
        order.brow.upload(brow.data(), brow.size(), st);
        order.bu.upload(bu.data(), bu.size(), st);
        order.ucols.upload(ucols.data(), ucols.size(), st);
        lcol.insert(lcol.end(), 8, (unsigned short)0);   // 16-byte chunk slack, as for the row pointers
        order.lcol.upload(lcol.data(), lcol.size(), st);
    }
    {
        std::vector<int> ofs, col, rslot;
        std::vector<rb_mask_t> mask;
        row_block_unions(n, prp, pci, ofs, col, mask, rslot);
        order.rb_nb = (int)ofs.size() - 1;
        // packed words of the pipelined kernel: row mask in the low 16 bits,
        // offset of the column's first coefficient within its block (column,
        // then row order) in the high 16 bits; 64 trailing entries so that a
        // batch's loads never leave the arrays
        std::vector<unsigned> word(col.size() + 64, 0u);
        bool fits = true;
        for (int b = 0; b < order.rb_nb && fits; b++) {
            unsigned off = 0;
            for (int u = ofs[b]; u < ofs[b + 1]; u++) {
                if (off > 0xffffu) { fits = false; break; }
                word[u] = (unsigned)mask[u] | (off << 16);
                off += (unsigned)__builtin_popcount((unsigned)mask[u]);
            }
        }
        std::vector<int> colp(col);
        colp.insert(colp.end(), 64, col.empty() ? 0 : col.back());
        order.rb_ofs.upload(ofs.data(), ofs.size(), st);
        order.rb_col.upload(colp.data(), colp.size(), st);
        order.rb_mask.upload(mask.data(), mask.size(), st);
        order.rb_slot.upload(rslot.data(), rslot.size(), st);
        order.rb_word_ok = fits;
        if (fits) order.rb_word.upload(word.data(), word.size(), st);
        TMI_HIP(hipStreamSynchronize(st));
    }
    TMI_HIP(hipStreamSynchronize(st));   // the host vectors go out of scope
    order.valid = true;
    return order;
}

}  // namespace tmi
