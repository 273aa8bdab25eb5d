// Wavefront OBJ ingest (host code): Mesh::load (mesh.cpp:186-263) reads a
// file through tinyobjloader's LoadObj with triangulation on and emits one
// TriangleGPU per face in file order, w = 1, _ModelId = the mesh's id.
//
// Subset read here (everything Mesh::load uses): `v x y z [w ...]` (missing
// coordinates are 0, extra values such as vertex colours ignored) and
// `f a b c ...` with `v`, `v/vt`, `v//vn`, `v/vt/vn` references, 1-based or
// negative (relative to the vertices read so far).  Other statements (vt, vn,
// g, o, s, usemtl, mtllib, l, p, comments) do not produce triangles.  Faces
// with fewer than 3 references are skipped, as tinyobjloader does.
// Triangulation: a quad is split along its shorter diagonal, (0,1,2)+(0,2,3)
// when |v2-v0|^2 < |v3-v1|^2, else (0,1,3)+(1,2,3), computed in float as
// tinyobjloader 2.x does; larger polygons become a fan (0,k,k+1).  Numbers are
// parsed with strtod and rounded to float.
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rtq/rtq.h"

namespace {

struct Parser {
    std::vector<float> v;                 // x, y, z per vertex
    std::vector<uint32_t> tri;            // 3 vertex indices per output triangle
    std::string err;
    uint64_t line = 0;

    bool fail(const char* what) {
        char buf[160];
        if (line)
            snprintf(buf, sizeof buf, "line %llu: %s", (unsigned long long)line, what);
        else
            snprintf(buf, sizeof buf, "%s", what);
        err = buf;
        return false;
    }

    static const char* skip_ws(const char* p, const char* e) {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
        return p;
    }

    static const char* token_end(const char* p, const char* e) {
        while (p < e && *p != ' ' && *p != '\t' && *p != '\r') ++p;
        return p;
    }

    bool vertex(const char* p, const char* e) {
        float xyz[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 3; ++k) {
            p = skip_ws(p, e);
            if (p == e) break;
            const char* q = token_end(p, e);
            std::string tok(p, q);
            char* end = nullptr;
            errno = 0;
            const double d = strtod(tok.c_str(), &end);
            if (end == tok.c_str() || *end != '\0') return fail("bad vertex coordinate");
            xyz[k] = (float)d;
            p = q;
        }
        v.insert(v.end(), xyz, xyz + 3);
        return true;
    }

    bool face(const char* p, const char* e) {
        std::vector<uint32_t> idx;
        const int64_t nv = (int64_t)(v.size() / 3);
        for (;;) {
            p = skip_ws(p, e);
            if (p == e) break;
            const char* q = token_end(p, e);
            // the vertex reference is the part before the first '/'
            const char* slash = (const char*)memchr(p, '/', (size_t)(q - p));
            std::string tok(p, slash ? slash : q);
            char* end = nullptr;
            const long long i = strtoll(tok.c_str(), &end, 10);
            if (tok.empty() || *end != '\0') return fail("bad face index");
            if (i == 0) return fail("face index 0");
            const int64_t a = i > 0 ? (int64_t)i - 1 : nv + (int64_t)i;
            if (a < 0) return fail("relative face index before the first vertex");
            if (a > 0xFFFFFFFFll) return fail("face index too large");
            idx.push_back((uint32_t)a);
            p = q;
        }
        const size_t n = idx.size();
        if (n < 3) return true;
        if (n == 3) {
            tri.insert(tri.end(), idx.begin(), idx.end());
        } else if (n == 4) {
            // checked against the vertex count once the file is read
            quads.push_back((uint64_t)tri.size());
            tri.insert(tri.end(), {idx[0], idx[1], idx[2], idx[0], idx[2], idx[3]});
            quad_alt.insert(quad_alt.end(), {idx[0], idx[1], idx[3], idx[1], idx[2], idx[3]});
        } else {
            for (size_t k = 1; k + 1 < n; ++k) tri.insert(tri.end(), {idx[0], idx[k], idx[k + 1]});
        }
        return true;
    }

    // quads: the diagonal needs the vertices, which may follow the face
    std::vector<uint64_t> quads;
    std::vector<uint32_t> quad_alt;

    void split_quads() {
        for (size_t q = 0; q < quads.size(); ++q) {
            uint32_t* t = tri.data() + quads[q];
            const uint32_t i0 = t[0], i1 = t[1], i2 = t[2], i3 = t[5];
            const float* v0 = &v[3ull * i0];
            const float* v1 = &v[3ull * i1];
            const float* v2 = &v[3ull * i2];
            const float* v3 = &v[3ull * i3];
            const float e02x = v2[0] - v0[0], e02y = v2[1] - v0[1], e02z = v2[2] - v0[2];
            const float e13x = v3[0] - v1[0], e13y = v3[1] - v1[1], e13z = v3[2] - v1[2];
            const float sqr02 = e02x * e02x + e02y * e02y + e02z * e02z;
            const float sqr13 = e13x * e13x + e13y * e13y + e13z * e13z;
            if (!(sqr02 < sqr13)) memcpy(t, &quad_alt[6 * q], 6 * sizeof(uint32_t));
        }
    }

    bool parse(const char* s, uint64_t len) {
        const char* p = s;
        const char* end = s + len;
        while (p < end) {
            const char* nl = (const char*)memchr(p, '\n', (size_t)(end - p));
            const char* e = nl ? nl : end;
            ++line;
            const char* q = skip_ws(p, e);
            if (q + 1 < e && q[0] == 'v' && (q[1] == ' ' || q[1] == '\t')) {
                if (!vertex(q + 2, e)) return false;
            } else if (q + 1 < e && q[0] == 'f' && (q[1] == ' ' || q[1] == '\t')) {
                if (!face(q + 2, e)) return false;
            }
            p = nl ? nl + 1 : end;
        }
        const uint64_t nv = v.size() / 3;
        for (uint32_t i : tri)
            if (i >= nv) {
                line = 0;
                return fail("face index past the last vertex");
            }
        for (uint32_t i : quad_alt)
            if (i >= nv) {
                line = 0;
                return fail("face index past the last vertex");
            }
        split_quads();
        return true;
    }
};

rtq_status finish(Parser& ps, bool ok, uint32_t model_id, rtq_triangle* tris, uint64_t capacity, uint64_t* n_tris,
                  char* msg, uint64_t msg_capacity) {
    auto say = [&](const std::string& m) {
        if (msg && msg_capacity) {
            const size_t k = std::min<size_t>(m.size(), (size_t)msg_capacity - 1);
            memcpy(msg, m.data(), k);
            msg[k] = '\0';
        }
    };
    if (!ok) {
        say(ps.err);
        return RTQ_EINVAL;
    }
    const uint64_t n = ps.tri.size() / 3;
    *n_tris = n;
    say("");
    if (!tris) return RTQ_OK;
    if (capacity < n) {
        say("triangle buffer too small");
        return RTQ_ECAPACITY;
    }
    for (uint64_t t = 0; t < n; ++t) {
        rtq_triangle& o = tris[t];
        memset(&o, 0, sizeof o);
        rtq_vec4* ps_out[3] = {&o.p0, &o.p1, &o.p2};
        for (int k = 0; k < 3; ++k) {
            const float* s = &ps.v[3ull * ps.tri[3 * t + k]];
            *ps_out[k] = rtq_vec4{s[0], s[1], s[2], 1.0f};
        }
        o.model_id = model_id;
    }
    return RTQ_OK;
}

}  // namespace

extern "C" {

rtq_status rtq_obj_parse(const char* text, uint64_t len, uint32_t model_id, rtq_triangle* tris, uint64_t capacity,
                         uint64_t* n_tris, char* msg, uint64_t msg_capacity) {
    if ((!text && len) || !n_tris) return RTQ_EINVAL;
    Parser ps;
    const bool ok = ps.parse(text ? text : "", len);
    return finish(ps, ok, model_id, tris, capacity, n_tris, msg, msg_capacity);
}

rtq_status rtq_obj_load(const char* path, uint32_t model_id, rtq_triangle* tris, uint64_t capacity, uint64_t* n_tris,
                        char* msg, uint64_t msg_capacity) {
    if (!path || !n_tris) return RTQ_EINVAL;
    FILE* f = fopen(path, "rb");
    if (!f) {
        Parser ps;
        ps.err = std::string("cannot open ") + path;
        return finish(ps, false, model_id, tris, capacity, n_tris, msg, msg_capacity);
    }
    std::string data;
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) data.append(buf, k);
    fclose(f);
    Parser ps;
    const bool ok = ps.parse(data.data(), data.size());
    return finish(ps, ok, model_id, tris, capacity, n_tris, msg, msg_capacity);
}

}  // extern "C"
