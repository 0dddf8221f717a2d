// Host-side panel geometry and topology for the MI355X solve path.
//
// Replaces the topology half of geometry::readTri (geometry.cpp:132-484) with
// flat arrays: the device never sees pointers, only the resolved SoA tables
// built here (panel order, Kutta folding table, CHTLS clusters, wake lines).
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace cpgpu {

struct Vec3 {
    double x, y, z;
};

// One triangular or quadrilateral panel after setGeom (panel.cpp:33-104).
struct PanelGeom {
    int nv = 3;
    int node[4] = {-1, -1, -1, -1};
    Vec3 v[4];
    Vec3 center, normal, bez;
    double area = 0, long_side = 0;
};

// panel::setGeom (panel.cpp:33-104): centre, area, normal, longSide from v[]
void panel_setup(PanelGeom& p);

struct WakeLineInfo {
    int upper, lower;  // body panel indices (reference order)
    Vec3 p1, p2;       // trailing-edge end points (wakeLine::setDimensions)
    double y;          // pMid.y, the sort key (wake::addWakeLine)
};

struct WakeInfo {
    std::vector<int> panels;        // indices into wake panel order
    std::vector<WakeLineInfo> lines;
    double yMin = 0, yMax = 0, x0 = 0, xf = 0, z0 = 0, zf = 0;
};

// Vortex-particle mode (geometry::readTri with the vortex-particle flag,
// geometry.cpp:267-397; calcTimeStep :1276-1304; createVPWakeSurfaces
// :1307-1342): the mesh has no prescribed wake; every trailing-edge edge gets
// a first-row buffer quad (the wake whose doublets fold into A) and a
// second-row quad behind it (the columns of C).
struct VpParams {
    double V = 1;   // freestream speed of the node projections and of dt
    double dt = 0;  // time step; <= 0: mean trailing-edge edge length / V
};

// The second buffer row, in trailing-edge edge order.
struct VpBuffer {
    double dt = 0;
    std::vector<PanelGeom> panels;
    std::vector<int> parent;                    // first-row (wake) panel
    std::vector<std::array<Vec3, 4>> pio;       // wakePanel::pointsInOrder
    // per edge of wakePanel::edgesInOrder (j = 1..3): the second-row panel
    // across it (-1: none) and whether this panel is the first to use it
    // in panel order (cpCaseVP::collapseBufferWake counts an edge once)
    std::vector<std::array<int, 4>> eio_nbr;
    std::vector<std::array<char, 4>> eio_first;
    std::vector<std::array<int, 2>> rad_nbr;    // across getEdges()[1], [3]
};

// Fully resolved geometry of one .tri mesh.
struct Geometry {
    int n_nodes = 0;
    std::vector<Vec3> nodes;
    std::vector<PanelGeom> body;   // reference body order (surfaces prepended)
    std::vector<PanelGeom> wake;   // reference wake order
    std::vector<int> body_tri, wake_tri;  // file triangle index
    // Kutta folding table per wake panel (wakePanel::interpPanels)
    std::vector<int> fold_idx;     // 4 per wake panel: U1, L1, U2, L2
    std::vector<double> fold_coef; // interpCoeff t
    // CHTLS clusters (bodyPanel::setCluster), 16 entries, -1 padded
    std::vector<int> cluster;
    std::vector<int> ts_order;
    std::vector<char> tip, upper, lower;
    // trailing-edge flags of the upper TE panel edges needed by pntInWake
    std::vector<WakeInfo> wakes;
    // per wake panel: TE flag and upper body panel (for wake::pntInWake)
    std::vector<char> wake_te;
    std::vector<int> wake_upper;
    // per body panel: edges flagged TE (node pairs), up to 3
    std::vector<int> body_te_edges;  // 6 per body panel (n1,n2)*3, -1 if not TE
    // vortex-particle mode: the wake above is the first buffer row (wake_tri
    // is -1 for its quads, which no file triangle backs)
    bool vp = false;
    VpBuffer w2;
};

// Builds the geometry from 0-based connectivity.  Returns false with a
// message for meshes the reference would reject or mis-handle.
bool build_geometry(int n_nodes, const double* xyz, int n_tris, const int32_t* tri, const int32_t* ids,
                    Geometry& out, std::string& err, const VpParams* vp = nullptr);

bool read_tri_file(const std::string& path, std::vector<double>& xyz, std::vector<int32_t>& tri,
                   std::vector<int32_t>& ids, std::string& err);

// Simple parallel loop over [0, n) on the host using std::thread.
void parallel_for(int n, int nthreads, const std::function<void(int, int)>& body);

}  // namespace cpgpu
