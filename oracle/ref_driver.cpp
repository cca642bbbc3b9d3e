// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A caller for the reference's own host code (mesh.cpp, bvh.cpp, bbox.cpp, camera.cpp), compiled
// together with those files where they lie under the reference checkout into oracle/_ref/.  It
// exposes the reference's loaded mesh, light CDF, camera basis and six-order BVH in the canonical
// link form of ppm_scene, so the tests can check the product's host layer against the reference
// itself.  The link table is read the way CGPURT::CreateTextures lays out its hit/miss tiles
// (gpurt.cpp:326-473): texel k of face f holds the links of the face-f node whose idBase is k.
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bvh.h"
#include "camera.h"
#include "mesh.h"

extern std::vector<TMaterial> mtls;  // mesh.cpp's material list (never cleared by the reference)

namespace {
TMesh* g_mesh = nullptr;
TBVH* g_bvh = nullptr;
TCamera g_cam;

void reset() {
    delete g_mesh;
    g_mesh = new TMesh();
    g_bvh = nullptr;  // TBVH owns raw arrays and has no destructor; leaked on purpose
    mtls.clear();
}
}  // namespace

extern "C" {

// TMesh::LoadOBJ from inside the OBJ's directory (the reference opens the MTL relative to the cwd).
int ref_load_obj(const char* dir, const char* file, float scale, float px, float py, float pz) {
    reset();
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) return -1;
    if (chdir(dir) != 0) return -1;
    std::vector<char> name(file, file + std::strlen(file) + 1);
    g_mesh->LoadOBJ(name.data(), TVector3(px, py, pz), scale);
    if (chdir(cwd) != 0) return -1;
    return (int)g_mesh->triangles.size();
}

// A mesh from arrays, through the reference's own CalculateBBox and PrepareLightSources.
int ref_mesh_from_arrays(int T, const float* pos, const float* nrm, const int* mat, int M, const int* brdf) {
    reset();
    g_mesh->materials.resize(M);
    for (int m = 0; m < M; m++) g_mesh->materials[m].brdf = brdf[m];
    g_mesh->triangles.resize(T);
    for (int i = 0; i < T; i++) {
        TTriangle& t = g_mesh->triangles[i];
        for (int q = 0; q < 3; q++) {
            t.positions[q] = TVector3(pos[i * 9 + q * 3], pos[i * 9 + q * 3 + 1], pos[i * 9 + q * 3 + 2]);
            t.normals[q] = TVector3(nrm[i * 9 + q * 3], nrm[i * 9 + q * 3 + 1], nrm[i * 9 + q * 3 + 2]);
            t.texcoords[q] = TVector2(1.0e+30f, 1.0e+30f);
        }
        t.idMaterial = mat[i];
    }
    g_mesh->CalculateBBox();
    bool light = false;
    for (int m = 0; m < M; m++) light = light || brdf[m] == -1;
    if (light) g_mesh->PrepareLightSources();
    else g_mesh->lightsArea = 0.0f;
    return T;
}

// Texture coordinates [T][3][2] and the textures of the loaded mesh (TMesh::LoadOBJ, LoadPPM).
void ref_texcoords(float* out) {
    for (size_t i = 0; i < g_mesh->triangles.size(); i++)
        for (int q = 0; q < 3; q++) {
            out[i * 6 + q * 2] = g_mesh->triangles[i].texcoords[q].x;
            out[i * 6 + q * 2 + 1] = g_mesh->triangles[i].texcoords[q].y;
        }
}

int ref_texture_info(int m, int* width, int* height) {
    const TMaterial& t = g_mesh->materials[m];
    *width = t.isTextured ? t.textureWidth : 0;
    *height = t.isTextured ? t.textureHeight : 0;
    return t.isTextured ? 1 : 0;
}

void ref_texture(int m, unsigned char* out) {
    const TMaterial& t = g_mesh->materials[m];
    if (t.isTextured) std::memcpy(out, t.texture.data(), t.texture.size());
}

int ref_num_triangles() { return g_mesh ? (int)g_mesh->triangles.size() : 0; }
int ref_num_materials() { return g_mesh ? (int)g_mesh->materials.size() : 0; }
int ref_num_lights() { return g_mesh ? (int)g_mesh->lightsCDF.size() : 0; }

void ref_triangles(float* pos, float* nrm, int* mat) {
    for (size_t i = 0; i < g_mesh->triangles.size(); i++) {
        const TTriangle& t = g_mesh->triangles[i];
        for (int q = 0; q < 3; q++) {
            pos[i * 9 + q * 3 + 0] = t.positions[q].x;
            pos[i * 9 + q * 3 + 1] = t.positions[q].y;
            pos[i * 9 + q * 3 + 2] = t.positions[q].z;
            nrm[i * 9 + q * 3 + 0] = t.normals[q].x;
            nrm[i * 9 + q * 3 + 1] = t.normals[q].y;
            nrm[i * 9 + q * 3 + 2] = t.normals[q].z;
        }
        mat[i] = t.idMaterial;
    }
}

// (r, g, b, eta, brdf, g) per material: the two texels of TextureMaterials (gpurt.cpp:184-198)
void ref_materials(float* out) {
    for (size_t i = 0; i < g_mesh->materials.size(); i++) {
        const TMaterial& m = g_mesh->materials[i];
        out[i * 6 + 0] = m.color.x;
        out[i * 6 + 1] = m.color.y;
        out[i * 6 + 2] = m.color.z;
        out[i * 6 + 3] = m.eta;
        out[i * 6 + 4] = (float)m.brdf;
        out[i * 6 + 5] = 1.0f / (pow(2.0f / (1.0f - m.specularity), 2.71828f) + 1.0f);
    }
}

float ref_lights(int* tri, float* cdf) {
    for (size_t i = 0; i < g_mesh->lightsCDF.size(); i++) {
        tri[i] = g_mesh->lightsIndices[i];
        cdf[i] = g_mesh->lightsCDF[i];
    }
    return g_mesh->lightsArea;
}

void ref_bbox(float* out) {
    out[0] = g_mesh->bbox.min.x;
    out[1] = g_mesh->bbox.min.y;
    out[2] = g_mesh->bbox.min.z;
    out[3] = g_mesh->bbox.max.x;
    out[4] = g_mesh->bbox.max.y;
    out[5] = g_mesh->bbox.max.z;
}

int ref_build_bvh() {
    g_bvh = new TBVH();
    g_bvh->Build(*g_mesh);
    return g_bvh->nodesNum;
}

// bounds[N][6], tri[N], links[6][N][2] in canonical node order
void ref_bvh(float* bounds, int* tri, int* links) {
    const int N = g_bvh->nodesNum;
    for (int k = 0; k < N; k++) {
        const TBVHNode& n = g_bvh->nodes[0][k];
        bounds[k * 6 + 0] = n.bbox.min.x;
        bounds[k * 6 + 1] = n.bbox.min.y;
        bounds[k * 6 + 2] = n.bbox.min.z;
        bounds[k * 6 + 3] = n.bbox.max.x;
        bounds[k * 6 + 4] = n.bbox.max.y;
        bounds[k * 6 + 5] = n.bbox.max.z;
        tri[k] = n.isLeaf ? n.idTriangle : -1;
    }
    for (int f = 0; f < 6; f++) {
        TBVHNode* nf = g_bvh->nodes[f];
        for (int i = 0; i < N; i++) {
            const int k = nf[i].idBase;
            const int hit = (i + 1 >= N) ? -1 : nf[i + 1].idBase;
            const int m = nf[i].idMiss;
            const int miss = (m == -1 || m >= N) ? -1 : nf[m].idBase;
            links[(f * N + k) * 2 + 0] = hit;
            links[(f * N + k) * 2 + 1] = miss;
        }
    }
}

// TCamera::Set; out = origin, u, v, w, half width, half height, distance
void ref_camera(const float* origin, const float* lookat, int w, int h, float fov, float* out) {
    g_cam.Set(TVector3(origin[0], origin[1], origin[2]), TVector3(lookat[0], lookat[1], lookat[2]), w, h, fov);
    const TVector3* v[4] = {&g_cam.origin, &g_cam.u, &g_cam.v, &g_cam.w};
    for (int k = 0; k < 4; k++) {
        out[k * 3 + 0] = v[k]->x;
        out[k * 3 + 1] = v[k]->y;
        out[k * 3 + 2] = v[k]->z;
    }
    out[12] = (float)g_cam.width * 0.5f;
    out[13] = (float)g_cam.height * 0.5f;
    out[14] = g_cam.distance;
}

}  // extern "C"
