// Host scene layer of the SPPM pass loop: meshes, materials, light CDF, camera, and the six-order
// threaded BVH flattened into ppm_scene.
//
// The float expressions here that feed the pass loop (bounding boxes, centroids, SAH costs,
// light areas, camera basis) keep the operation order of the reference's host code so that the
// tree and the uniforms come out with the same bits:
//   TBBox::Expand/Area (bbox.cpp:4-34), TMesh::CalculateBBox/PrepareLightSources
//   (mesh.cpp:437-488), TCamera::Set (camera.cpp:4-18), splitBVH/sortAxis (bvh.cpp:7-174),
//   the face re-orderings and miss links (bvh.cpp:178-317).
// Build with strict IEEE float (no contraction, no fast-math).

#include "../../include/ppm.h"

#include <algorithm>
#include <cctype>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_mesh_error;

int fail(int code, const std::string& msg) {
    g_mesh_error = msg;
    return code;
}

struct V3 {
    float x, y, z;
};
inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 muls(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// TVector3::normalize: v * (1 / sqrt(dot)), float sqrt (vector.h:12).
inline V3 unit(V3 a) { return muls(a, 1.0f / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z)); }
inline float get(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

struct Box {
    V3 lo, hi;
    void reset() {
        lo = mk(1e+38f, 1e+38f, 1e+38f);
        hi = mk(-1e+38f, -1e+38f, -1e+38f);
    }
    // max keeps the current value unless the point is larger (vector.h Maximize/Minimize).
    void grow(V3 p) {
        hi = mk(hi.x > p.x ? hi.x : p.x, hi.y > p.y ? hi.y : p.y, hi.z > p.z ? hi.z : p.z);
        lo = mk(lo.x < p.x ? lo.x : p.x, lo.y < p.y ? lo.y : p.y, lo.z < p.z ? lo.z : p.z);
    }
    float area() const {
        V3 l = sub(hi, lo);
        return 2.0f * (l.x * l.y + l.y * l.z + l.z * l.x);
    }
};

struct Tri {
    V3 p[3];
    V3 n[3];
    int mat;
    float uv[3][2] = {{1.0e30f, 1.0e30f}, {1.0e30f, 1.0e30f}, {1.0e30f, 1.0e30f}};  // as given
};

struct Node {
    Box box;
    int leaf;   // 1 for leaves
    int tri;    // leaf triangle or -1
    int left, right;
};

}  // namespace

struct ppm_mesh {
    std::vector<Tri> tris;
    std::vector<ppm_material> mats;
    ppm_camera cam;
    bool has_camera = false;
    bool has_uv = false;                        // every triangle carries texture coordinates
    std::vector<std::vector<uint8_t>> tex;      // per material: RGB8 image, empty if untextured
    std::vector<int> tex_w, tex_h;

    // derived by build
    Box bbox;
    std::vector<float> positions, normals;
    std::vector<int32_t> material;
    std::vector<float> node_bounds;
    std::vector<int32_t> node_triangle, node_links;
    std::vector<int32_t> light_triangle;
    std::vector<float> light_cdf;
    float light_area = 0.0f;
    std::vector<float> texcoords;
    std::vector<uint8_t> volume;
    ppm_scene scene;
    bool built = false;
};

namespace {

// ------------------------------------------------------------------------------------------
// Threaded multi-order BVH.
//
// Face 0 is an exact-SAH tree with one triangle per leaf, numbered in pre-order; the child with
// the smaller min.x goes first (bvh.cpp:68-210).  The sweep sorts triangle indices by centroid
// with the same in-place partition as sortAxis (bvh.cpp:7-29) because the order of equal keys
// decides the sweep and so the tree.  Faces 1-5 reorder children by max.x, min.y, max.y, min.z,
// max.z (bvh.cpp:292-296).  In every order the hit link is the next node in pre-order and the
// miss link is the first node after the subtree (bvh.cpp:211-241, gpurt.cpp:332-361).
// ------------------------------------------------------------------------------------------

// Same partition steps as sortAxis, with an explicit stack that visits the left range to the end
// before the right one, as the recursion does.
void sort_by_key(int* idx, const float* key, int lo, int hi, std::vector<std::pair<int, int>>& st) {
    st.clear();
    st.push_back({lo, hi});
    while (!st.empty()) {
        int li = st.back().first, ri = st.back().second;
        st.pop_back();
        int i = li, j = ri;
        const float pivot = key[idx[(li + ri) / 2]];
        for (;;) {
            while (key[idx[i]] < pivot) i++;
            while (key[idx[j]] > pivot) j--;
            if (i >= j) break;
            int t = idx[i];
            idx[i] = idx[j];
            idx[j] = t;
            i++;
            j--;
        }
        // right range pushed first so the left range is finished first
        if ((j + 1) < ri) st.push_back({j + 1, ri});
        if (li < (i - 1)) st.push_back({li, i - 1});
    }
}

// sort_by_key for large ranges: after each partition step the two ranges are disjoint and sorted
// independently, so the right one goes to another thread; the permutation is the same.
void sort_by_key_par(int* idx, const float* key, int lo, int hi, int depth) {
    if (depth <= 0 || hi - lo < 131072) {
        std::vector<std::pair<int, int>> st;
        sort_by_key(idx, key, lo, hi, st);
        return;
    }
    int i = lo, j = hi;
    const float pivot = key[idx[(lo + hi) / 2]];
    for (;;) {
        while (key[idx[i]] < pivot) i++;
        while (key[idx[j]] > pivot) j--;
        if (i >= j) break;
        int t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
        i++;
        j--;
    }
    std::thread right;
    if ((j + 1) < hi) right = std::thread(sort_by_key_par, idx, key, j + 1, hi, depth - 1);
    if (lo < (i - 1)) sort_by_key_par(idx, key, lo, i - 1, depth - 1);
    if (right.joinable()) right.join();
}

struct BuildTask {
    int lo, n;       // range into the index pool
    Box box;         // box handed down by the parent's sweep
    int parent;      // -1 for the root
    int side;        // 0 = first child in the canonical order, 1 = second
};

// Per-thread buffers of the SAH sweep; sort_depth > 0 lets the node's sorts use extra threads.
struct Scratch {
    int sort_depth = 0;
    std::vector<int> sorted;
    std::vector<float> left_area;
    std::vector<Box> lbox;
    std::vector<std::pair<int, int>> sstack;
    void fit(int n) {
        if ((int)sorted.size() < n) {
            sorted.resize(n);
            left_area.resize(n);
            lbox.resize(n);
        }
    }
};

struct BuildInput {
    std::vector<Box> tbox;
    std::vector<float> cen[3];
    std::vector<int> pool;
};

// One node of splitBVH (bvh.cpp:68-210): an exact SAH sweep over the three axes, each axis sorting
// the range as the previous axis left it (the order of equal keys matters), then the range
// reordered as the best axis sorted it.  Returns false for a leaf; otherwise the two children in
// the canonical order (the side with the smaller min.x first, bvh.cpp:195).
bool split_node(BuildInput& in, Scratch& sc, const BuildTask& tk, int id, BuildTask& first, BuildTask& second) {
    if (tk.n <= 1) return false;
    int* idx = &in.pool[tk.lo];
    const int n = tk.n;
    int best_index = 0;
    Box best_l, best_r;
    best_l.lo = best_l.hi = mk(0.0f, 0.0f, 0.0f);
    best_r = best_l;
    int nl, nr;
    if (n == 2) {
        best_l = in.tbox[idx[0]];
        best_r = in.tbox[idx[1]];
        nl = 1;
        nr = 1;
    } else {
        sc.fit(n);
        int best_axis = 0;
        float best_cost = 1e+30f;
        for (int k = 0; k <= 2; k++) {
            if (sc.sort_depth > 0 && n >= 262144) sort_by_key_par(idx, in.cen[k].data(), 0, n - 1, sc.sort_depth);
            else sort_by_key(idx, in.cen[k].data(), 0, n - 1, sc.sstack);
            Box bl;
            bl.reset();
            for (int i = 0; i < n; i++) {
                bl.grow(in.tbox[idx[i]].lo);
                bl.grow(in.tbox[idx[i]].hi);
                sc.left_area[i] = bl.area();
                sc.lbox[i] = bl;
            }
            int tri_num = n - 1;
            Box br;
            br.reset();
            for (int j = n - 2; j >= 0; j--) {
                const int ii = idx[j + 1];
                br.grow(in.tbox[ii].lo);
                br.grow(in.tbox[ii].hi);
                const float cost = (float)tri_num * sc.left_area[j] + (float)(n - tri_num) * br.area();
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = k;
                    best_index = j;
                    best_l = sc.lbox[best_index];
                    best_r = br;
                }
                tri_num--;
            }
            if (best_axis == k) std::memcpy(sc.sorted.data(), idx, sizeof(int) * n);
        }
        std::memcpy(idx, sc.sorted.data(), sizeof(int) * n);
        nl = best_index + 1;
        nr = n - nl;
    }
    BuildTask lt = {tk.lo, nl, best_l, id, 0};
    BuildTask rt = {tk.lo + nl, nr, best_r, id, 0};
    if (best_l.lo.x < best_r.lo.x) {
        first = lt;
        second = rt;
    } else {
        first = rt;
        second = lt;
    }
    first.side = 0;
    second.side = 1;
    return true;
}

// A node as the parallel build records it: ids are handed out in completion order and turned
// into the canonical pre-order at the end.
struct RawNode {
    Box box;
    int tri;            // leaf triangle, -1 for an internal node
    int child[2];       // canonical first and second child
};

int build_threads() {
    const char* e = std::getenv("PPM_BUILD_THREADS");
    int n = e ? std::atoi(e) : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

// Builds the subtree of `root` depth first on the calling thread (ranges below the cutoff).
void build_serial(BuildInput& in, Scratch& sc, const BuildTask& root, std::vector<RawNode>& raw,
                  std::atomic<int>& next_id) {
    std::vector<BuildTask> stack;
    stack.push_back(root);
    while (!stack.empty()) {
        BuildTask tk = stack.back();
        stack.pop_back();
        const int id = next_id.fetch_add(1);
        RawNode& r = raw[id];
        r.box = tk.box;
        r.child[0] = r.child[1] = -1;
        if (tk.parent >= 0) raw[tk.parent].child[tk.side] = id;
        BuildTask a, b;
        if (!split_node(in, sc, tk, id, a, b)) {
            r.tri = tk.n ? in.pool[tk.lo] : -1;
            continue;
        }
        r.tri = -1;
        stack.push_back(b);
        stack.push_back(a);
    }
}

// The canonical tree (face 0 pre-order ids, left = first child) of splitBVH.  Every node's split
// depends only on its own range of the index pool, as its parent left it, and on its box, so
// disjoint subtrees are built on separate threads: large ranges are split by a pool of workers,
// ranges below a cutoff are finished depth first by one worker.  The result does not depend on
// the number of threads.
void build_canonical(const std::vector<Tri>& tris, const Box& root_box, std::vector<Node>& nodes) {
    const int T = (int)tris.size();
    BuildInput in;
    in.tbox.resize(T);
    for (int k = 0; k < 3; k++) in.cen[k].resize(T);
    for (int i = 0; i < T; i++) {
        Box b;
        b.reset();
        b.grow(tris[i].p[0]);
        b.grow(tris[i].p[1]);
        b.grow(tris[i].p[2]);
        in.tbox[i] = b;
        V3 c = muls(add(add(tris[i].p[0], tris[i].p[1]), tris[i].p[2]), 1.0f / 3.0f);
        in.cen[0][i] = c.x;
        in.cen[1][i] = c.y;
        in.cen[2][i] = c.z;
    }
    in.pool.resize(T);
    for (int i = 0; i < T; i++) in.pool[i] = i;
    const int total = T > 0 ? 2 * T - 1 : 1;
    std::vector<RawNode> raw(total);
    std::atomic<int> next_id(0);

    const int threads = build_threads();
    const int cutoff = std::max(1024, T / (8 * threads));
    if (threads == 1 || T <= cutoff) {
        Scratch sc;
        build_serial(in, sc, BuildTask{0, T, root_box, -1, 0}, raw, next_id);
    } else {
        std::mutex mu;
        std::condition_variable cv;
        std::vector<BuildTask> queue;
        queue.push_back(BuildTask{0, T, root_box, -1, 0});
        int busy = 0;
        int depth = 0;
        while ((2 << depth) <= threads) depth++;
        auto worker = [&]() {
            Scratch sc;
            sc.sort_depth = depth;
            for (;;) {
                BuildTask tk;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return !queue.empty() || busy == 0; });
                    if (queue.empty()) return;
                    // the largest waiting range first keeps the critical path short
                    auto it = std::max_element(queue.begin(), queue.end(),
                                               [](const BuildTask& x, const BuildTask& y) { return x.n < y.n; });
                    tk = *it;
                    queue.erase(it);
                    busy++;
                }
                if (tk.n <= cutoff) {
                    build_serial(in, sc, tk, raw, next_id);
                } else {
                    const int id = next_id.fetch_add(1);
                    RawNode& r = raw[id];
                    r.box = tk.box;
                    r.tri = -1;
                    r.child[0] = r.child[1] = -1;
                    BuildTask a, b;
                    const bool inner = split_node(in, sc, tk, id, a, b);
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        if (tk.parent >= 0) raw[tk.parent].child[tk.side] = id;
                        if (inner) {
                            queue.push_back(a);
                            queue.push_back(b);
                        } else {
                            r.tri = tk.n ? in.pool[tk.lo] : -1;
                        }
                    }
                }
                {
                    std::lock_guard<std::mutex> lk(mu);
                    busy--;
                }
                cv.notify_all();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++) pool.emplace_back(worker);
        for (auto& th : pool) th.join();
    }

    // canonical pre-order: node, first child's subtree, second child's subtree
    const int N = next_id.load();
    nodes.assign(N, Node());
    std::vector<int> order;
    order.reserve(N);
    std::vector<int> st;
    if (N > 0) st.push_back(0);
    std::vector<int> canon(N, -1);
    while (!st.empty()) {
        const int k = st.back();
        st.pop_back();
        canon[k] = (int)order.size();
        order.push_back(k);
        if (raw[k].child[0] >= 0) {
            st.push_back(raw[k].child[1]);
            st.push_back(raw[k].child[0]);
        }
    }
    for (int c = 0; c < N; c++) {
        const RawNode& r = raw[order[c]];
        Node& nd = nodes[c];
        nd.box = r.box;
        nd.leaf = r.child[0] < 0 ? 1 : 0;
        nd.tri = nd.leaf ? r.tri : -1;
        nd.left = nd.leaf ? -1 : canon[r.child[0]];
        nd.right = nd.leaf ? -1 : canon[r.child[1]];
    }
}

// True when face f keeps the canonical child order at an internal node (bvh.cpp:328-332).
bool keep_order(int face, const Node& a, const Node& b) {
    switch (face) {
        case 0: return true;
        case 1: return a.box.hi.x > b.box.hi.x;
        case 2: return a.box.lo.y < b.box.lo.y;
        case 3: return a.box.hi.y > b.box.hi.y;
        case 4: return a.box.lo.z < b.box.lo.z;
        default: return a.box.hi.z > b.box.hi.z;
    }
}

void thread_links(const std::vector<Node>& nodes, std::vector<int32_t>& links) {
    const int N = (int)nodes.size();
    links.assign((size_t)6 * N * 2, -1);
    // subtree sizes; children always have larger canonical ids than their parent
    std::vector<int> size(N, 1);
    for (int k = N - 1; k >= 0; k--)
        if (!nodes[k].leaf) size[k] = 1 + size[nodes[k].left] + size[nodes[k].right];
    // the six orders are independent: one thread each for large trees
    auto face_links = [&](int f) {
        std::vector<int> order(N), pos(N), st;
        int c = 0;
        if (N > 0) st.push_back(0);
        while (!st.empty()) {
            int k = st.back();
            st.pop_back();
            order[c] = k;
            pos[k] = c;
            c++;
            if (nodes[k].leaf) continue;
            int a = nodes[k].left, b = nodes[k].right;
            if (!keep_order(f, nodes[a], nodes[b])) std::swap(a, b);
            st.push_back(b);
            st.push_back(a);
        }
        for (int k = 0; k < N; k++) {
            const int p = pos[k];
            const int hit = (p + 1 < N) ? order[p + 1] : -1;
            const int mp = p + size[k];
            const int miss = (k == 0 || mp >= N) ? -1 : order[mp];
            links[((size_t)f * N + k) * 2 + 0] = hit;
            links[((size_t)f * N + k) * 2 + 1] = miss;
        }
    };
    if (N < 65536 || build_threads() == 1) {
        for (int f = 0; f < 6; f++) face_links(f);
    } else {
        std::vector<std::thread> th;
        for (int f = 0; f < 6; f++) th.emplace_back(face_links, f);
        for (auto& t : th) t.join();
    }
}

// gpurt.cpp:194, evaluated with float pow as the reference's host code does.
float gloss_of(float specularity) {
    return 1.0f / (std::pow(2.0f / (1.0f - specularity), 2.71828f) + 1.0f);
}

// ------------------------------------------------------------------------------------------
// Seeded scene generators.
// ------------------------------------------------------------------------------------------

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform float in [a, b)
    float uni(float a, float b) {
        double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
        return (float)(a + (b - a) * u);
    }
};

int add_material(ppm_mesh* m, float r, float g, float b, int brdf, float eta) {
    ppm_material mt;
    mt.color[0] = r;
    mt.color[1] = g;
    mt.color[2] = b;
    mt.eta = eta;
    mt.brdf = brdf;
    mt.g = gloss_of(1.0f);
    m->mats.push_back(mt);
    return (int)m->mats.size() - 1;
}

// Flat-shaded triangle; the geometric normal cross(p1-p0, p2-p0) is the facing side.
void add_flat(ppm_mesh* m, V3 a, V3 b, V3 c, int mat) {
    Tri t;
    t.p[0] = a;
    t.p[1] = b;
    t.p[2] = c;
    V3 n = unit(cross3(sub(b, a), sub(c, a)));
    t.n[0] = t.n[1] = t.n[2] = n;
    t.mat = mat;
    m->tris.push_back(t);
}

void add_quad(ppm_mesh* m, V3 a, V3 b, V3 c, V3 d, int mat) {
    add_flat(m, a, b, c, mat);
    add_flat(m, a, c, d, mat);
}

// Axis-aligned box resting on y = y0, rotated about the vertical axis; faces point outwards.
void add_box(ppm_mesh* m, float cx, float cz, float sx, float sy, float sz, float angle, float y0,
             int mat) {
    const float ca = std::cos(angle), sa = std::sin(angle);
    V3 c[8];
    for (int i = 0; i < 8; i++) {
        float x = ((i & 1) ? 0.5f : -0.5f) * sx;
        float y = (i & 2) ? sy : 0.0f;
        float z = ((i & 4) ? 0.5f : -0.5f) * sz;
        c[i] = mk(cx + ca * x + sa * z, y0 + y, cz - sa * x + ca * z);
    }
    // (corner order chosen so that each face's normal points away from the box centre)
    add_quad(m, c[0], c[1], c[5], c[4], mat);  // bottom (-y)
    add_quad(m, c[2], c[6], c[7], c[3], mat);  // top (+y)
    add_quad(m, c[0], c[4], c[6], c[2], mat);  // -x
    add_quad(m, c[1], c[3], c[7], c[5], mat);  // +x
    add_quad(m, c[0], c[2], c[3], c[1], mat);  // -z
    add_quad(m, c[4], c[5], c[7], c[6], mat);  // +z
}

// Unit room [-0.5, 0.5]^3 open towards +z, a 0.25 x 0.25 ceiling light, a tall mirror box and a
// short diffuse box: 36 triangles, the make-up of cornell_metal.obj.
void gen_cornell(ppm_mesh* m) {
    const int white = add_material(m, 0.75f, 0.75f, 0.75f, PPM_BRDF_DIFFUSE, 1.7f);
    const int red = add_material(m, 0.75f, 0.25f, 0.25f, PPM_BRDF_DIFFUSE, 1.7f);
    const int green = add_material(m, 0.25f, 0.75f, 0.25f, PPM_BRDF_DIFFUSE, 1.7f);
    const int light = add_material(m, 1.0f, 1.0f, 1.0f, PPM_BRDF_LIGHT, 1.7f);
    const int metal = add_material(m, 0.8f, 0.8f, 0.8f, PPM_BRDF_MIRROR, 1.7f);
    const float h = 0.5f;
    // inward-facing walls
    add_quad(m, mk(-h, -h, -h), mk(-h, -h, h), mk(h, -h, h), mk(h, -h, -h), white);   // floor +y
    add_quad(m, mk(-h, h, -h), mk(h, h, -h), mk(h, h, h), mk(-h, h, h), white);       // ceiling -y
    add_quad(m, mk(-h, -h, -h), mk(h, -h, -h), mk(h, h, -h), mk(-h, h, -h), white);   // back +z
    // the look-at basis has u = up x w = -x for a camera on +z, so +x is the viewer's left
    add_quad(m, mk(-h, -h, -h), mk(-h, h, -h), mk(-h, h, h), mk(-h, -h, h), green);   // -x wall
    add_quad(m, mk(h, -h, -h), mk(h, -h, h), mk(h, h, h), mk(h, h, -h), red);         // +x wall
    // light quad just under the ceiling, facing down
    const float l = 0.125f, ly = h - 0.0025f;
    add_quad(m, mk(-l, ly, -l), mk(l, ly, -l), mk(l, ly, l), mk(-l, ly, l), light);
    add_box(m, -0.17f, -0.14f, 0.3f, 0.6f, 0.3f, 0.35f, -h, metal);
    add_box(m, 0.18f, 0.12f, 0.3f, 0.3f, 0.3f, -0.30f, -h, white);
}

void add_icosphere(ppm_mesh* m, SplitMix& rng, V3 c, float r, int detail, float jitter, int mat) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    std::vector<V3> v = {mk(-1, t, 0), mk(1, t, 0), mk(-1, -t, 0), mk(1, -t, 0),
                         mk(0, -1, t), mk(0, 1, t), mk(0, -1, -t), mk(0, 1, -t),
                         mk(t, 0, -1), mk(t, 0, 1), mk(-t, 0, -1), mk(-t, 0, 1)};
    for (auto& p : v) p = unit(p);
    std::vector<int> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                          11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                          3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < detail; s++) {
        std::map<std::pair<int, int>, int> mid;
        std::vector<int> nf;
        nf.reserve(f.size() * 4);
        auto midpoint = [&](int a, int b) {
            std::pair<int, int> key(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V3 p = unit(muls(add(v[a], v[b]), 0.5f));
            v.push_back(p);
            int id = (int)v.size() - 1;
            mid[key] = id;
            return id;
        };
        for (size_t k = 0; k < f.size(); k += 3) {
            int a = f[k], b = f[k + 1], cc = f[k + 2];
            int ab = midpoint(a, b), bc = midpoint(b, cc), ca = midpoint(cc, a);
            int add4[12] = {a, ab, ca, b, bc, ab, cc, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), add4, add4 + 12);
        }
        f.swap(nf);
    }
    std::vector<V3> wp(v.size()), wn(v.size());
    for (size_t i = 0; i < v.size(); i++) {
        V3 p = muls(v[i], r);
        V3 j = mk(rng.uni(-jitter, jitter), rng.uni(-jitter, jitter), rng.uni(-jitter, jitter));
        p = add(p, muls(j, r));
        wp[i] = add(c, p);
        wn[i] = unit(p);
    }
    for (size_t k = 0; k < f.size(); k += 3) {
        Tri tr;
        // icosahedron faces above are counter-clockwise seen from outside
        for (int q = 0; q < 3; q++) {
            tr.p[q] = wp[f[k + q]];
            tr.n[q] = wn[f[k + q]];
        }
        tr.mat = mat;
        m->tris.push_back(tr);
    }
}

// Cornell room plus `count` icospheres: 50% diffuse (albedo 0.2-0.8 per channel), 30% mirror,
// 20% glass with IOR 1.5; radii 0.03-0.12, centres uniform inside the room, 1% vertex jitter.
void gen_spheres(ppm_mesh* m, uint64_t seed, int count, int detail) {
    gen_cornell(m);
    SplitMix rng{seed};
    for (int s = 0; s < count; s++) {
        const float r = rng.uni(0.03f, 0.12f);
        V3 c = mk(rng.uni(-0.5f + r, 0.5f - r), rng.uni(-0.5f + r, 0.5f - r), rng.uni(-0.5f + r, 0.5f - r));
        const float pick = rng.uni(0.0f, 1.0f);
        int mat;
        if (pick < 0.5f) {
            float a = rng.uni(0.2f, 0.8f), b = rng.uni(0.2f, 0.8f), d = rng.uni(0.2f, 0.8f);
            mat = add_material(m, a, b, d, PPM_BRDF_DIFFUSE, 1.7f);
        } else if (pick < 0.8f) {
            float a = rng.uni(0.6f, 0.95f);
            mat = add_material(m, a, a, a, PPM_BRDF_MIRROR, 1.7f);
        } else {
            mat = add_material(m, 1.0f, 1.0f, 1.0f, PPM_BRDF_GLASS, 1.5f);
        }
        add_icosphere(m, rng, c, r, detail, 0.01f, mat);
    }
}

// Value noise on an integer lattice, smooth-step interpolated.
float lattice(uint64_t seed, int x, int z) {
    uint64_t h = seed ^ ((uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)z * 0xC2B2AE3D27D4EB4Full);
    SplitMix s{h};
    return (float)((s.next() >> 40) * (1.0 / 16777216.0));
}
float value_noise(uint64_t seed, float x, float z) {
    int xi = (int)std::floor(x), zi = (int)std::floor(z);
    float fx = x - xi, fz = z - zi;
    float sx = fx * fx * (3.0f - 2.0f * fx), sz = fz * fz * (3.0f - 2.0f * fz);
    float a = lattice(seed, xi, zi), b = lattice(seed, xi + 1, zi);
    float c = lattice(seed, xi, zi + 1), d = lattice(seed, xi + 1, zi + 1);
    return (a + (b - a) * sx) * (1.0f - sz) + (c + (d - c) * sx) * sz;
}

// detail x detail heightfield over [-1, 1]^2 displaced by 5 octaves of value noise, `count`
// glass icospheres of subdivision `sphere_detail`, one small area light above.
void gen_terrain(ppm_mesh* m, uint64_t seed, int count, int detail, int sphere_detail) {
    const int ground = add_material(m, 0.6f, 0.55f, 0.45f, PPM_BRDF_DIFFUSE, 1.7f);
    const int light = add_material(m, 1.0f, 1.0f, 1.0f, PPM_BRDF_LIGHT, 1.7f);
    const int glass = add_material(m, 1.0f, 1.0f, 1.0f, PPM_BRDF_GLASS, 1.5f);
    const int n = detail;
    std::vector<V3> p((size_t)(n + 1) * (n + 1));
    for (int j = 0; j <= n; j++)
        for (int i = 0; i <= n; i++) {
            float x = -1.0f + 2.0f * i / n, z = -1.0f + 2.0f * j / n;
            float hgt = 0.0f, amp = 0.5f, fr = 4.0f;
            for (int o = 0; o < 5; o++) {
                hgt += amp * value_noise(seed + o, (x + 1.0f) * fr, (z + 1.0f) * fr);
                amp *= 0.5f;
                fr *= 2.0f;
            }
            p[(size_t)j * (n + 1) + i] = mk(x, 0.25f * hgt - 0.4f, z);
        }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            V3 a = p[(size_t)j * (n + 1) + i], b = p[(size_t)j * (n + 1) + i + 1];
            V3 c = p[(size_t)(j + 1) * (n + 1) + i + 1], d = p[(size_t)(j + 1) * (n + 1) + i];
            // upward facing: (a, d, c), (a, c, b)
            add_flat(m, a, d, c, ground);
            add_flat(m, a, c, b, ground);
        }
    const float l = 0.1f, ly = 0.9f;
    add_quad(m, mk(-l, ly, -l), mk(l, ly, -l), mk(l, ly, l), mk(-l, ly, l), light);
    SplitMix rng{seed * 31 + 7};
    for (int s = 0; s < count; s++) {
        float r = rng.uni(0.08f, 0.16f);
        V3 c = mk(rng.uni(-0.7f, 0.7f), rng.uni(0.0f, 0.3f), rng.uni(-0.7f, 0.7f));
        add_icosphere(m, rng, c, r, sphere_detail, 0.0f, glass);
    }
}

// ------------------------------------------------------------------------------------------
// OBJ / MTL (mesh.cpp:37-427, 491-693).  Lines are read whole; a material is committed by its
// Ns line; names keep their line terminator as in the reference's comparisons.
// ------------------------------------------------------------------------------------------

struct MtlRec {
    std::string name;
    V3 ka{0, 0, 0}, kd{0, 0, 0}, ks{0, 0, 0};
    float ns = 0.0f;
    bool textured = false;
    int tw = 0, th = 0;
    std::vector<uint8_t> tex;
};

// LoadPPM (mesh.cpp:37-55): "P6", width, height, "255", each followed by white space as
// fscanf's "\n" skips it (so are leading white-space bytes of the pixel data, as there), then
// width x height x 3 bytes; a short file leaves the rest zero.  Comments are not read.
bool load_ppm_image(const std::string& path, int& w, int& h, std::vector<uint8_t>& rgb) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    std::vector<uint8_t> d;
    int ch;
    while ((ch = std::fgetc(fp)) != EOF) d.push_back((uint8_t)ch);
    std::fclose(fp);
    size_t i = 0;
    auto skip_ws = [&]() {
        while (i < d.size() && std::isspace(d[i])) i++;
    };
    auto read_int = [&](int& v) -> bool {
        skip_ws();
        size_t j = i;
        long long x = 0;
        while (i < d.size() && d[i] >= '0' && d[i] <= '9' && x < (1ll << 31)) x = x * 10 + (d[i++] - '0');
        v = (int)x;
        return i > j;
    };
    if (d.size() < 2 || d[0] != 'P' || d[1] != '6') return false;
    i = 2;
    skip_ws();
    int maxv = 0;
    if (!read_int(w) || !read_int(h) || !read_int(maxv) || maxv != 255 || w <= 0 || h <= 0 ||
        (long long)w * h > (1ll << 28))
        return false;
    skip_ws();
    rgb.assign((size_t)w * h * 3, 0);
    const size_t n = std::min(rgb.size(), d.size() - i);
    if (n) std::memcpy(rgb.data(), d.data() + i, n);
    return true;
}

bool starts(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

std::string dir_of(const std::string& path) {
    size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

bool read_lines(const std::string& path, std::vector<std::string>& out) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    std::string cur;
    int ch;
    while ((ch = std::fgetc(fp)) != EOF) {
        cur.push_back((char)ch);
        if (ch == '\n') {
            out.push_back(cur);
            cur.clear();
        }
    }
    if (!cur.empty()) out.push_back(cur);
    std::fclose(fp);
    return true;
}

bool load_mtl(const std::string& path, std::vector<MtlRec>& mtls) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return false;
    MtlRec cur;
    for (const std::string& ln : lines) {
        float r = 0, g = 0, b = 0, s = 0;
        if (starts(ln, "newmtl")) {
            cur.name = ln.size() > 7 ? ln.substr(7) : std::string();
            cur.textured = false;
        } else if (starts(ln, "Ka") || starts(ln, "Ke")) {
            std::sscanf(ln.c_str() + 3, "%f %f %f", &r, &g, &b);
            cur.ka = mk(r, g, b);
        } else if (starts(ln, "Kd")) {
            std::sscanf(ln.c_str() + 3, "%f %f %f", &r, &g, &b);
            cur.kd = mk(r, g, b);
        } else if (starts(ln, "Ks")) {
            std::sscanf(ln.c_str() + 3, "%f %f %f", &r, &g, &b);
            cur.ks = mk(r, g, b);
        } else if (starts(ln, "Ns")) {
            std::sscanf(ln.c_str() + 3, "%f", &s);
            cur.ns = s;
            mtls.push_back(cur);
            mtls.back().textured = false;
            mtls.back().tex.clear();
        } else if (starts(ln, "map_Kd")) {
            // the image belongs to the last material committed so far (mesh.cpp:117-122);
            // the name is resolved next to the MTL file first, then as given
            std::string name = ln.size() > 7 ? ln.substr(7) : std::string();
            while (!name.empty() && (name.back() == '\n' || name.back() == '\r')) name.pop_back();
            if (mtls.empty()) continue;
            MtlRec& m = mtls.back();
            if (!load_ppm_image(dir_of(path) + name, m.tw, m.th, m.tex) &&
                !load_ppm_image(name, m.tw, m.th, m.tex))
                return false;
            m.textured = true;
        }
    }
    return true;
}

// mesh.cpp:520-611
ppm_material classify(const MtlRec& r) {
    ppm_material m;
    int brdf = PPM_BRDF_DIFFUSE;
    float eta = 1.7f, spec = 1.0f;
    const float ks2 = dot3(r.ks, r.ks);
    if (r.ns == 100.0f) {
        if (ks2 == 3.0f) brdf = PPM_BRDF_MIRROR;
        else if (ks2 > 0.0f) brdf = PPM_BRDF_PLASTIC;
    } else {
        spec = std::max(r.ns / 100.0f, 0.5f);
        if (ks2 == 3.0f) brdf = PPM_BRDF_GLOSSY_MIRROR;
        else if (ks2 > 0.0f) brdf = PPM_BRDF_GLOSSY_PLASTIC;
    }
    if (starts(r.name, "glass")) {
        eta = ks2 / 3.0f + 1.0f;
        if (r.ns == 100.0f) {
            brdf = PPM_BRDF_GLASS;
        } else {
            spec = std::max(r.ns / 100.0f, 0.5f);
            brdf = PPM_BRDF_GLOSSY_GLASS;
        }
    }
    if (dot3(r.ka, r.ka) > 0.0f) brdf = PPM_BRDF_LIGHT;
    V3 kd = r.kd;
    if (brdf == PPM_BRDF_DIFFUSE || brdf == PPM_BRDF_PLASTIC) kd = muls(kd, 0.9f);
    if (brdf == PPM_BRDF_GLASS || brdf == PPM_BRDF_GLOSSY_GLASS)
        kd = mk(std::sqrt(kd.x), std::sqrt(kd.y), std::sqrt(kd.z));
    if (starts(r.name, "sss")) brdf = 7;  // no shader branch handles it (kept for fidelity)
    m.color[0] = kd.x;
    m.color[1] = kd.y;
    m.color[2] = kd.z;
    m.eta = eta;
    m.brdf = brdf;
    m.g = gloss_of(spec);
    return m;
}

void parse_corner(const char* w, int& vi, int& ti, int& ni) {
    // "v", "v/t", "v//n", "v/t/n"; missing entries read as 0 (atoi of an empty field)
    vi = ti = ni = 0;
    vi = std::atoi(w);
    const char* s1 = std::strchr(w, '/');
    if (!s1) return;
    ti = std::atoi(s1 + 1);
    const char* s2 = std::strchr(s1 + 1, '/');
    if (!s2) return;
    ni = std::atoi(s2 + 1);
}

int load_obj(const char* path, const float position[3], float scale, ppm_mesh* m) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return fail(PPM_E_IO, std::string("cannot read ") + path);
    std::vector<MtlRec> mtls;
    std::vector<V3> v, vn;
    std::vector<float> vt;
    struct Face {
        int v[3], t[3], n[3];
        int mat;
    };
    std::vector<Face> faces;
    bool no_normals = false, no_materials = true, no_texcoords = false;
    int cmat = 0;
    for (const std::string& ln : lines) {
        if (starts(ln, "mtllib")) {
            std::string name = ln.size() > 7 ? ln.substr(7) : std::string();
            while (!name.empty() && (name.back() == '\n' || name.back() == '\r')) name.pop_back();
            if (!load_mtl(dir_of(path) + name, mtls) && !load_mtl(name, mtls))
                return fail(PPM_E_IO, "cannot read material library " + name);
        }
        if (ln.size() > 1 && ln[0] == 'v') {
            float x = 0, y = 0, z = 0;
            if (ln[1] == 'n') {
                std::sscanf(ln.c_str() + 2, "%f %f %f", &x, &y, &z);
                float l = std::sqrt(x * x + y * y + z * z);
                vn.push_back(mk(x / l, y / l, z / l));
            } else if (ln[1] == 't') {
                float u = 0, w = 0;
                std::sscanf(ln.c_str() + 2, "%f %f", &u, &w);
                vt.push_back(u);
                vt.push_back(w);
            } else {
                std::sscanf(ln.c_str() + 1, "%f %f %f", &x, &y, &z);
                v.push_back(mk(x, y, z));
            }
        }
        if (starts(ln, "usemtl")) {
            std::string name = ln.size() > 7 ? ln.substr(7) : std::string();
            for (size_t i = 0; i < mtls.size(); i++)
                if (name == mtls[i].name) {
                    cmat = (int)i;
                    no_materials = false;
                    break;
                }
        } else if (!ln.empty() && ln[0] == 'f') {
            char s1[64] = {0}, s2[64] = {0}, s3[64] = {0};
            std::sscanf(ln.c_str() + 1, "%63s %63s %63s", s1, s2, s3);
            Face fc;
            const char* w[3] = {s1, s2, s3};
            for (int q = 0; q < 3; q++) {
                int vi, ti, ni;
                parse_corner(w[q], vi, ti, ni);
                fc.v[q] = vi - 1;
                fc.t[q] = ti - 1;
                fc.n[q] = ni - 1;
                if (!ni) no_normals = true;
                if (!ti) no_texcoords = true;
            }
            fc.mat = cmat;
            faces.push_back(fc);
        }
    }
    const V3 pos = mk(position[0], position[1], position[2]);
    if (!no_materials) {
        for (const MtlRec& r : mtls) {
            m->mats.push_back(classify(r));
            m->tex.push_back(r.textured ? r.tex : std::vector<uint8_t>());
            m->tex_w.push_back(r.textured ? r.tw : 0);
            m->tex_h.push_back(r.textured ? r.th : 0);
        }
    } else {
        // mesh.cpp:613-621: a default material whose colour is left at zero
        ppm_material d;
        d.color[0] = d.color[1] = d.color[2] = 0.0f;
        d.eta = 1.0f;
        d.brdf = PPM_BRDF_DIFFUSE;
        d.g = 0.0f;
        m->mats.push_back(d);
        m->tex.emplace_back();
        m->tex_w.push_back(0);
        m->tex_h.push_back(0);
    }
    m->has_uv = !no_texcoords && !faces.empty();
    for (const Face& fc : faces) {
        Tri t;
        for (int q = 0; q < 3; q++) {
            if (fc.v[q] < 0 || fc.v[q] >= (int)v.size()) return fail(PPM_E_INVAL, "bad vertex index");
            t.p[q] = add(muls(v[fc.v[q]], scale), pos);
        }
        if (!no_normals) {
            for (int q = 0; q < 3; q++) {
                if (fc.n[q] < 0 || fc.n[q] >= (int)vn.size()) return fail(PPM_E_INVAL, "bad normal index");
                t.n[q] = vn[fc.n[q]];
            }
        } else {
            V3 n = unit(cross3(sub(t.p[1], t.p[0]), sub(t.p[2], t.p[0])));
            t.n[0] = t.n[1] = t.n[2] = n;
        }
        t.mat = no_materials ? 0 : fc.mat;
        for (int q = 0; q < 3; q++) {
            t.uv[q][0] = t.uv[q][1] = 1.0e30f;
            if (m->has_uv) {
                if (fc.t[q] < 0 || 2 * (size_t)fc.t[q] + 1 >= vt.size())
                    return fail(PPM_E_INVAL, "bad texture coordinate index");
                t.uv[q][0] = vt[2 * (size_t)fc.t[q]];
                t.uv[q][1] = vt[2 * (size_t)fc.t[q] + 1];
            }
        }
        m->tris.push_back(t);
    }
    return PPM_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------

extern "C" {

const char* ppm_mesh_last_error(void) { return g_mesh_error.c_str(); }

int ppm_mesh_load_obj(const char* path, const float position[3], float scale, ppm_mesh** out) {
    if (!path || !out) return fail(PPM_E_INVAL, "null argument");
    const float zero[3] = {0, 0, 0};
    ppm_mesh* m = new (std::nothrow) ppm_mesh();
    if (!m) return fail(PPM_E_OOM, "out of memory");
    int rc = load_obj(path, position ? position : zero, scale, m);
    if (rc != PPM_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return PPM_OK;
}

int ppm_mesh_from_arrays(int32_t num_triangles, const float* positions, const float* normals,
                         const int32_t* material, int32_t num_materials,
                         const ppm_material* materials, ppm_mesh** out) {
    if (!out || num_triangles < 0 || (num_triangles > 0 && !positions) || num_materials < 1 ||
        !materials)
        return fail(PPM_E_INVAL, "bad mesh arrays");
    ppm_mesh* m = new (std::nothrow) ppm_mesh();
    if (!m) return fail(PPM_E_OOM, "out of memory");
    m->mats.assign(materials, materials + num_materials);
    m->tris.resize(num_triangles);
    for (int i = 0; i < num_triangles; i++) {
        Tri& t = m->tris[i];
        for (int q = 0; q < 3; q++) {
            const float* p = positions + (size_t)i * 9 + q * 3;
            t.p[q] = mk(p[0], p[1], p[2]);
        }
        if (normals) {
            for (int q = 0; q < 3; q++) {
                const float* p = normals + (size_t)i * 9 + q * 3;
                t.n[q] = mk(p[0], p[1], p[2]);
            }
        } else {
            V3 n = unit(cross3(sub(t.p[1], t.p[0]), sub(t.p[2], t.p[0])));
            t.n[0] = t.n[1] = t.n[2] = n;
        }
        t.mat = material ? material[i] : 0;
        if (t.mat < 0 || t.mat >= num_materials) {
            delete m;
            return fail(PPM_E_INVAL, "material index out of range");
        }
    }
    *out = m;
    return PPM_OK;
}

int ppm_mesh_set_texcoords(ppm_mesh* m, const float* texcoords) {
    if (!m) return fail(PPM_E_INVAL, "null mesh");
    for (size_t i = 0; i < m->tris.size(); i++)
        for (int q = 0; q < 3; q++) {
            m->tris[i].uv[q][0] = texcoords ? texcoords[i * 6 + q * 2] : 1.0e30f;
            m->tris[i].uv[q][1] = texcoords ? texcoords[i * 6 + q * 2 + 1] : 1.0e30f;
        }
    m->has_uv = texcoords != nullptr;
    m->built = false;
    return PPM_OK;
}

int ppm_mesh_set_texture(ppm_mesh* m, int32_t material, int32_t width, int32_t height, const uint8_t* rgb) {
    if (!m) return fail(PPM_E_INVAL, "null mesh");
    if (material < 0 || material >= (int)m->mats.size()) return fail(PPM_E_INVAL, "material out of range");
    if (rgb && (width <= 0 || height <= 0 || (long long)width * height > (1ll << 28)))
        return fail(PPM_E_INVAL, "bad texture size");
    m->tex.resize(m->mats.size());
    m->tex_w.resize(m->mats.size(), 0);
    m->tex_h.resize(m->mats.size(), 0);
    if (rgb) m->tex[material].assign(rgb, rgb + (size_t)width * height * 3);
    else m->tex[material].clear();
    m->tex_w[material] = rgb ? width : 0;
    m->tex_h[material] = rgb ? height : 0;
    m->built = false;
    return PPM_OK;
}

int ppm_mesh_generate(int32_t kind, uint64_t seed, int32_t count, int32_t detail, ppm_mesh** out) {
    if (!out) return fail(PPM_E_INVAL, "null argument");
    if (count < 0 || detail < 0 || detail > 8192) return fail(PPM_E_INVAL, "bad generator sizes");
    ppm_mesh* m = new (std::nothrow) ppm_mesh();
    if (!m) return fail(PPM_E_OOM, "out of memory");
    try {
        if (kind == PPM_SCENE_CORNELL) {
            gen_cornell(m);
        } else if (kind == PPM_SCENE_SPHERES) {
            if (detail > 9) {
                delete m;
                return fail(PPM_E_INVAL, "icosphere subdivision above 9");
            }
            gen_spheres(m, seed, count, detail);
        } else if (kind == PPM_SCENE_TERRAIN) {
            gen_terrain(m, seed, count, detail > 0 ? detail : 64, 5);
        } else {
            delete m;
            return fail(PPM_E_INVAL, "unknown scene kind");
        }
    } catch (const std::bad_alloc&) {
        delete m;
        return fail(PPM_E_OOM, "out of memory");
    }
    *out = m;
    return PPM_OK;
}

int ppm_mesh_set_camera(ppm_mesh* m, const float origin[3], const float lookat[3], int32_t width,
                        int32_t height, float fov) {
    if (!m || !origin || !lookat || width <= 0 || height <= 0) return fail(PPM_E_INVAL, "bad camera");
    ppm_camera& c = m->cam;
    const V3 o = mk(origin[0], origin[1], origin[2]);
    const V3 at = mk(lookat[0], lookat[1], lookat[2]);
    const float dist = float(height) / (2.0f * std::tan((fov / 2.0f) * (3.141592f / 180.0f)));
    const V3 tv = mk(0.0f, 1.0f, 0.0f);
    V3 w = unit(sub(at, o));
    V3 u = unit(cross3(tv, w));
    V3 v = unit(cross3(w, u));
    for (int k = 0; k < 3; k++) {
        c.origin[k] = get(o, k);
        c.u[k] = get(u, k);
        c.v[k] = get(v, k);
        c.w[k] = get(w, k);
    }
    c.half_width = (float)width * 0.5f;
    c.half_height = (float)height * 0.5f;
    c.distance = dist;
    m->has_camera = true;
    if (m->built) m->scene.camera = c;
    return PPM_OK;
}

int ppm_mesh_build(ppm_mesh* m) {
    if (!m) return fail(PPM_E_INVAL, "null mesh");
    if (m->tris.empty()) return fail(PPM_E_INVAL, "mesh has no triangles");
    try {
        const int T = (int)m->tris.size();
        // TMesh::CalculateBBox
        m->bbox.reset();
        for (const Tri& t : m->tris)
            for (int q = 0; q < 3; q++) m->bbox.grow(t.p[q]);
        // TMesh::PrepareLightSources
        m->light_triangle.clear();
        m->light_cdf.clear();
        m->light_area = 0.0f;
        for (int i = 0; i < T; i++) {
            const Tri& t = m->tris[i];
            if (m->mats[t.mat].brdf != PPM_BRDF_LIGHT) continue;
            V3 c = cross3(sub(t.p[1], t.p[0]), sub(t.p[2], t.p[0]));
            const float a = 0.5f * std::sqrt(dot3(c, c));
            m->light_cdf.push_back(a);
            m->light_triangle.push_back(i);
            m->light_area += a;
        }
        for (size_t i = 1; i < m->light_cdf.size(); i++) m->light_cdf[i] = m->light_cdf[i] + m->light_cdf[i - 1];
        for (size_t i = 0; i < m->light_cdf.size(); i++) m->light_cdf[i] = m->light_cdf[i] / m->light_area;

        std::vector<Node> nodes;
        build_canonical(m->tris, m->bbox, nodes);
        const int N = (int)nodes.size();
        m->node_bounds.resize((size_t)N * 6);
        m->node_triangle.resize(N);
        for (int k = 0; k < N; k++) {
            const Box& b = nodes[k].box;
            float* o = &m->node_bounds[(size_t)k * 6];
            o[0] = b.lo.x; o[1] = b.lo.y; o[2] = b.lo.z;
            o[3] = b.hi.x; o[4] = b.hi.y; o[5] = b.hi.z;
            m->node_triangle[k] = nodes[k].leaf ? nodes[k].tri : -1;
        }
        thread_links(nodes, m->node_links);

        // TMesh::LoadOBJ (mesh.cpp:656-669): a triangle keeps its texture coordinates only when
        // its material is textured; then the volume of gpurt.cpp:158-198 (nearest resampling to
        // 512^2, magenta for untextured materials)
        const int M = (int)m->mats.size();
        m->tex.resize(M);
        m->tex_w.resize(M, 0);
        m->tex_h.resize(M, 0);
        bool any_tex = false;
        m->texcoords.assign((size_t)T * 6, 1.0e30f);
        if (m->has_uv)
            for (int i = 0; i < T; i++)
                if (!m->tex[m->tris[i].mat].empty()) {
                    any_tex = true;
                    for (int q = 0; q < 3; q++) {
                        m->texcoords[(size_t)i * 6 + q * 2] = m->tris[i].uv[q][0];
                        m->texcoords[(size_t)i * 6 + q * 2 + 1] = m->tris[i].uv[q][1];
                    }
                }
        const int S = 512;
        m->volume.clear();
        if (any_tex) {
            m->volume.resize((size_t)M * S * S * 3);
            for (int z = 0; z < M; z++)
                for (int y = 0; y < S; y++)
                    for (int x = 0; x < S; x++) {
                        uint8_t* d = &m->volume[3 * ((size_t)x + (size_t)y * S + (size_t)z * S * S)];
                        if (!m->tex[z].empty()) {
                            const int tw = m->tex_w[z], th = m->tex_h[z];
                            const size_t src = 3 * ((size_t)(x * tw / S) + (size_t)(y * th / S) * tw);
                            d[0] = m->tex[z][src];
                            d[1] = m->tex[z][src + 1];
                            d[2] = m->tex[z][src + 2];
                        } else {
                            d[0] = 255;
                            d[1] = 0;
                            d[2] = 255;
                        }
                    }
        } else {
            m->texcoords.clear();
        }

        m->positions.resize((size_t)T * 9);
        m->normals.resize((size_t)T * 9);
        m->material.resize(T);
        for (int i = 0; i < T; i++) {
            for (int q = 0; q < 3; q++) {
                m->positions[(size_t)i * 9 + q * 3 + 0] = m->tris[i].p[q].x;
                m->positions[(size_t)i * 9 + q * 3 + 1] = m->tris[i].p[q].y;
                m->positions[(size_t)i * 9 + q * 3 + 2] = m->tris[i].p[q].z;
                m->normals[(size_t)i * 9 + q * 3 + 0] = m->tris[i].n[q].x;
                m->normals[(size_t)i * 9 + q * 3 + 1] = m->tris[i].n[q].y;
                m->normals[(size_t)i * 9 + q * 3 + 2] = m->tris[i].n[q].z;
            }
            m->material[i] = m->tris[i].mat;
        }
        ppm_scene& s = m->scene;
        std::memset(&s, 0, sizeof(s));
        s.num_triangles = T;
        s.positions = m->positions.data();
        s.normals = m->normals.data();
        s.material = m->material.data();
        s.num_materials = (int32_t)m->mats.size();
        s.materials = m->mats.data();
        s.num_nodes = N;
        s.node_bounds = m->node_bounds.data();
        s.node_triangle = m->node_triangle.data();
        s.node_links = m->node_links.data();
        s.num_lights = (int32_t)m->light_triangle.size();
        s.light_triangle = m->light_triangle.empty() ? nullptr : m->light_triangle.data();
        s.light_cdf = m->light_cdf.empty() ? nullptr : m->light_cdf.data();
        s.light_area = m->light_area;
        // SceneBSphere (glslppm.cpp:388-392)
        s.bsphere[0] = (m->bbox.hi.x + m->bbox.lo.x) * 0.5f;
        s.bsphere[1] = (m->bbox.hi.y + m->bbox.lo.y) * 0.5f;
        s.bsphere[2] = (m->bbox.hi.z + m->bbox.lo.z) * 0.5f;
        V3 d = sub(m->bbox.hi, m->bbox.lo);
        s.bsphere[3] = std::sqrt(dot3(d, d)) * 0.5f;
        if (!m->has_camera) {
            const float o[3] = {0.0f, 0.0f, 1.667f}, at[3] = {0.0f, 0.0f, 0.0f};
            ppm_mesh_set_camera(m, o, at, 256, 256, 45.0f);
        }
        s.camera = m->cam;
        s.texcoords = m->texcoords.empty() ? nullptr : m->texcoords.data();
        s.texture_size = m->volume.empty() ? 0 : 512;
        s.textures = m->volume.empty() ? nullptr : m->volume.data();
        m->built = true;
    } catch (const std::bad_alloc&) {
        return fail(PPM_E_OOM, "out of memory building the BVH");
    }
    return PPM_OK;
}

const ppm_scene* ppm_mesh_scene(const ppm_mesh* m) { return (m && m->built) ? &m->scene : nullptr; }

int32_t ppm_mesh_num_triangles(const ppm_mesh* m) { return m ? (int32_t)m->tris.size() : 0; }

void ppm_mesh_destroy(ppm_mesh* m) { delete m; }

}  // extern "C"
