// Internal structures shared by the host analysis (symbolic.cpp, numeric.cpp)
// and the device runtime (fcw_api.hip). Not part of the public C-ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fcw {

// Separator-tree supernodal structure of the cotan stiffness pattern.
//
// Vertices are renumbered by a nested-dissection ordering ("positions").
// Every tree node owns a contiguous position range [start, start + s) (its
// separator, or a whole leaf subdomain) and a sorted list of b boundary
// positions > start + s (the ancestor vertices adjacent to its subtree).
// Its factor panel is dense (s + b) x s, column-major.
struct Symbolic {
    int n = 0;                    // vertices
    int m = 0;                    // faces
    std::vector<int> perm;        // perm[p] = input vertex at position p
    std::vector<int> iperm;       // iperm[v] = position of input vertex v

    int nnodes = 0;
    std::vector<int> node_start, node_s, node_b, node_parent, node_level;
    std::vector<int64_t> node_boff;       // offset into bidx / cmap
    std::vector<int> bidx;                // boundary positions, per node, ascending
    std::vector<int> cmap;                // per node: its boundary rows -> parent front rows
    std::vector<int64_t> node_poff;       // panel offset (elements)
    std::vector<int64_t> node_uoff;       // update-buffer row offset
    std::vector<int> child_ptr, child_list;
    int nlevels = 0;
    std::vector<int> level_ptr, level_nodes;  // nodes grouped by level (leaves = level 0)
    int64_t panel_elems = 0;
    int64_t urows = 0;
    // Children's update rows feeding each row, two slots per row (-1 = none; the separator tree is
    // binary): contrib_own[2p + i] for position p, contrib_tail[2u + i] for update row u.
    std::vector<int> contrib_own, contrib_tail;

    // Full symmetric CSR of the operator pattern in positions (sorted columns, with diagonal).
    std::vector<int> rowptr, colind, diag;
    // Per CSR entry: contributing (face*9 + a*3 + b) with faces in position numbering.
    std::vector<int> contrib_ptr, contrib;
    // Faces in position numbering, reordered for locality.
    std::vector<int> faces;               // m x 3
    std::vector<int> face_perm;           // face_perm[f_new] = input face index
    // Vertex -> incident (face*3 + corner), position numbering.
    std::vector<int> vf_ptr, vf;

    // Work counts.
    double factor_flops = 0.0;       // one numeric factorisation (incl. inverse panels)
    double solve_flops_per_rhs = 0.0;  // one forward + one backward sweep, one RHS column
    int64_t factor_nnz = 0;          // sum over nodes of s(s+1)/2 + s*b (Cholesky factor entries)
    int max_front = 0;
};

// par_depth: dissection levels whose two halves are built on their own threads (the tree and the
// numbering do not depend on it); < 0 selects 2 on meshes of 20,000 vertices or more, else 0.
int analyse(int n, const double* V, int m, const int* F, int leaf_size, Symbolic& S, std::string& err,
            int par_depth = -1);

// Intrinsic Delaunay triangulation of the surface (V, F) after mollification with
// eps = mollify * mean edge length (intrinsic.cpp): same vertices, flipped faces, and the length
// of the edge opposite each corner.
// A mesh with non-manifold edges (or neighbours of inconsistent orientation) is flipped on its
// tufted cover, whose faces come back merged with their mirror copies: weight 1 for a merged pair,
// 1/2 for a cover face left without its copy (the operators use area * weight).
struct Intrinsic {
    std::vector<int> faces;      // mo x 3 (mo = m for an edge-manifold mesh, at most 2m on the cover)
    std::vector<double> len;     // mo x 3: edge opposite corner c
    std::vector<double> weight;  // mo
    int flips = 0;
    int skipped = 0;             // non-Delaunay edges whose flip would fold a face
    double delta = 0.0;          // added to every edge length
    int cover = 0;               // 1: built on the tufted cover
    int nonmanifold_edges = 0;   // input edges with three or more faces
};
constexpr double kMollify = 1e-6;
int intrinsic_delaunay(int n, const double* V, int m, const int* F, double mollify, Intrinsic& out, std::string& err);
// Area and basis-gradient Gram matrix (G00 G01 G02 G11 G12 G22) of each face from its lengths.
void intrinsic_operators(int m, const double* len, double* area, double* gram);

// Multifrontal FP64 Cholesky of A = diag_scale*Mdiag + off_scale*L (pattern of S, values in CSR
// order `lval`, lumped mass `mdiag` in positions). Writes the partitioned-inverse panels
// P_S = [L11^{-1}; L21 L11^{-1}] into `panel` (S.panel_elems doubles).
int factor(const Symbolic& S, const double* lval, const double* mdiag, double l_scale, double m_scale,
           double* panel, std::string& err, int nthreads);

// CPU application of the panels: solves A X = B in place for k columns (row-major n x k, positions).
void panel_solve_host(const Symbolic& S, const double* panel, double* X, int k);

}  // namespace fcw
