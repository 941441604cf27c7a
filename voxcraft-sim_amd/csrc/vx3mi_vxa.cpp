// vx3mi_vxa.cpp -- VXA/VXD input: a small XML reader, the VXD merge rule and the
// VXA semantics of the reference's model construction path.
//
//   merge      ctool::ptree_merge (src/Utils/ctool.h:49-57): every child of
//              <VXD> with a replace="VXA.a.b" attribute replaces the subtree at
//              that path of the base VXA (parents are created when missing)
//   simulator  CVX_Sim::ReadXML (src/VXA/VX_Sim.cpp:185-294) and the VX3 keys
//              read by VX3_SimulationManager::readVXD (VX3_SimulationManager.cu:323-376)
//   environment CVX_Environment::ReadXML (src/old/VX_Environment.cpp:101-171)
//   palette    CVXC_Material::ReadXML (src/VXA/VX_Object.cpp:1373-1475)
//   structure  CVXC_Structure::ReadXML (src/VXA/VX_Object.cpp:1760-1958)
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "vx3mi_internal.h"

namespace vx3mi {
namespace {

struct XNode {
    std::string name, text;
    bool cdata = false;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XNode> kids;
    const XNode *child(const std::string &n) const {
        for (const XNode &k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
    const char *attr(const std::string &n) const {
        for (const auto &a : attrs)
            if (a.first == n) return a.second.c_str();
        return nullptr;
    }
};

struct XmlError {
    std::string msg;
};

class XmlReader {
  public:
    explicit XmlReader(const char *s) : p_(s) {}
    XNode parse_document() {
        XNode doc;
        doc.name = "";
        skip_misc();
        while (*p_) {
            if (*p_ != '<') throw XmlError{"text outside the root element"};
            doc.kids.push_back(element());
            skip_misc();
        }
        if (doc.kids.empty()) throw XmlError{"no root element"};
        return doc;
    }

  private:
    const char *p_;

    static bool name_char(char c) { return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':'; }
    void ws() {
        while (*p_ && std::isspace((unsigned char)*p_)) p_++;
    }
    bool starts(const char *t) const { return std::strncmp(p_, t, std::strlen(t)) == 0; }
    void skip_until(const char *t) {
        const char *e = std::strstr(p_, t);
        if (!e) throw XmlError{std::string("unterminated construct, expected ") + t};
        p_ = e + std::strlen(t);
    }
    void skip_misc() {
        for (;;) {
            ws();
            if (starts("<?")) skip_until("?>");
            else if (starts("<!--")) skip_until("-->");
            else if (starts("<!DOCTYPE")) skip_until(">");
            else return;
        }
    }
    std::string name() {
        const char *b = p_;
        while (*p_ && name_char(*p_)) p_++;
        if (p_ == b) throw XmlError{"expected a name"};
        return std::string(b, p_);
    }
    static void decode_into(const char *b, const char *e, std::string &out) {
        while (b < e) {
            if (*b == '&') {
                const char *semi = (const char *)std::memchr(b, ';', e - b);
                if (!semi) { out.push_back(*b++); continue; }
                std::string ent(b + 1, semi);
                if (ent == "lt") out.push_back('<');
                else if (ent == "gt") out.push_back('>');
                else if (ent == "amp") out.push_back('&');
                else if (ent == "quot") out.push_back('"');
                else if (ent == "apos") out.push_back('\'');
                else if (!ent.empty() && ent[0] == '#') {
                    long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                                : std::strtol(ent.c_str() + 1, nullptr, 10);
                    out.push_back((char)c);
                } else {
                    out.append(b, semi + 1);
                }
                b = semi + 1;
            } else {
                out.push_back(*b++);
            }
        }
    }
    XNode element() {
        XNode n;
        p_++;  // '<'
        n.name = name();
        for (;;) {
            ws();
            if (*p_ == '/') {
                if (p_[1] != '>') throw XmlError{"bad empty-element tag"};
                p_ += 2;
                return n;
            }
            if (*p_ == '>') { p_++; break; }
            std::string an = name();
            ws();
            if (*p_ != '=') throw XmlError{"attribute without value"};
            p_++;
            ws();
            char q = *p_;
            if (q != '"' && q != '\'') throw XmlError{"unquoted attribute"};
            const char *b = ++p_;
            while (*p_ && *p_ != q) p_++;
            if (!*p_) throw XmlError{"unterminated attribute"};
            std::string v;
            decode_into(b, p_, v);
            p_++;
            n.attrs.emplace_back(an, v);
        }
        std::string text;
        for (;;) {
            if (!*p_) throw XmlError{"unterminated element " + n.name};
            if (starts("<![CDATA[")) {
                p_ += 9;
                const char *e = std::strstr(p_, "]]>");
                if (!e) throw XmlError{"unterminated CDATA"};
                text.append(p_, e);
                n.cdata = true;
                p_ = e + 3;
            } else if (starts("<!--")) {
                skip_until("-->");
            } else if (starts("<?")) {
                skip_until("?>");
            } else if (starts("</")) {
                p_ += 2;
                std::string close = name();
                if (close != n.name) throw XmlError{"mismatched </" + close + "> for <" + n.name + ">"};
                ws();
                if (*p_ != '>') throw XmlError{"bad end tag"};
                p_++;
                break;
            } else if (*p_ == '<') {
                n.kids.push_back(element());
            } else {
                const char *b = p_;
                while (*p_ && *p_ != '<') p_++;
                decode_into(b, p_, text);
            }
        }
        if (n.cdata) n.text = text;
        else {
            size_t a = 0, e = text.size();
            while (a < e && std::isspace((unsigned char)text[a])) a++;
            while (e > a && std::isspace((unsigned char)text[e - 1])) e--;
            n.text = text.substr(a, e - a);
        }
        return n;
    }
};

// ctool::ptree_merge: put the VXD child at the dotted path of the base tree
void merge_vxd(XNode &doc, const XNode &vxdDoc) {
    const XNode *vxd = vxdDoc.child("VXD");
    if (!vxd) vxd = &vxdDoc;
    for (const XNode &k : vxd->kids) {
        const char *rep = k.attr("replace");
        if (!rep || !*rep) continue;
        std::string path(rep);
        XNode *cur = &doc;
        size_t b = 0;
        std::string last;
        for (;;) {
            size_t e = path.find('.', b);
            std::string comp = path.substr(b, e == std::string::npos ? std::string::npos : e - b);
            XNode *next = nullptr;
            for (XNode &c : cur->kids)
                if (c.name == comp) { next = &c; break; }
            if (!next) {
                cur->kids.emplace_back();
                cur->kids.back().name = comp;
                next = &cur->kids.back();
            }
            if (e == std::string::npos) {
                XNode repl = k;
                repl.name = comp;
                *next = repl;
                break;
            }
            cur = next;
            b = e + 1;
        }
    }
}

const XNode *path(const XNode *n, std::initializer_list<const char *> comps) {
    for (const char *c : comps) {
        if (!n) return nullptr;
        n = n->child(c);
    }
    return n;
}

// FindLoadElement conversions (src/old/XML_Rip.h:73-77): atof / atoi, bool = atoi != 0
bool load(const XNode *parent, const char *key, double &out) {
    const XNode *n = parent ? parent->child(key) : nullptr;
    if (!n) return false;
    out = std::atof(n->text.c_str());
    return true;
}
bool load(const XNode *parent, const char *key, float &out) {
    double d;
    if (!load(parent, key, d)) return false;
    out = (float)d;
    return true;
}
bool load(const XNode *parent, const char *key, int &out) {
    const XNode *n = parent ? parent->child(key) : nullptr;
    if (!n) return false;
    out = std::atoi(n->text.c_str());
    return true;
}

std::string base64_decode(const std::string &in) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    std::string out;
    int buf = 0, bits = 0;
    for (char c : in) {
        int v = val(c);
        if (v < 0) continue;
        buf = (buf << 6) | v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((char)((buf >> bits) & 0xff));
        }
    }
    return out;
}

std::vector<std::string> split_commas(const std::string &s) {
    std::vector<std::string> out;
    size_t b = 0;
    for (;;) {
        size_t e = s.find(',', b);
        out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
        if (e == std::string::npos) break;
        b = e + 1;
    }
    return out;
}

// One boundary-condition region (CVX_FRegion, src/old/VX_FRegion.cpp:165-219):
// a box, cylinder, sphere or triangle mesh in workspace fractions, the degrees
// of freedom it fixes and the displacement / force it applies.
struct BcRegion {
    int prim = 0;  // PRIM_BOX 0, PRIM_CYLINDER 1, PRIM_SPHERE 2, PRIM_MESH 3
    double X = 0, Y = 0, Z = 0, dX = 0, dY = 0, dZ = 0, radius = 0;
    std::vector<double> mv;  // mesh vertices (x, y, z each)
    std::vector<int> mf;     // mesh facets (three vertex indices each)
    double bbMin[3] = {0, 0, 0}, bbSize[3] = {0, 0, 0};
    int dof = 0;
    double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0}, disp[3] = {0, 0, 0}, angDisp[3] = {0, 0, 0};
};

struct VxaContent {
    vx3mi_lattice lat{};
    std::vector<BcRegion> bcs;
    double latticeDim = 0.001;              // Lattice_Dim
    double dimAdj[3] = {1.0, 1.0, 1.0};     // X/Y/Z_Dim_Adj
    std::vector<signed char> cells;
    std::vector<double> phase;
    std::vector<vx3mi_material> mats;
    std::vector<MtToken> stop, fitness;
    std::vector<MtToken> forceField[3];
    std::vector<MtToken> attachCond[5];
    std::string rawPrint;
    int savePositions = 0;
    int recordStepSize = 0, recordVoxel = 1, recordLink = 0;
    int secondaryExperiment = 0;
    double reinitAfter = 0.0;
    std::vector<double> ciliaBase, ciliaShift;  // 3 per voxel, voxel order
};

// CVX_FRegion::ReadXML (src/old/VX_FRegion.cpp:165-219)
int read_region(const XNode *fr, BcRegion &b, std::string &err) {
    if (!load(fr, "PrimType", b.prim)) { err = "FRegion without PrimType"; return VX3MI_E_PARSE; }
    if (b.prim < 0 || b.prim > 3) { err = "boundary-condition region PrimType must be 0..3"; return VX3MI_E_PARSE; }
    load(fr, "X", b.X); load(fr, "Y", b.Y); load(fr, "Z", b.Z);
    load(fr, "dX", b.dX); load(fr, "dY", b.dY); load(fr, "dZ", b.dZ);
    load(fr, "Radius", b.radius);
    int v;
    if (load(fr, "DofFixed", v)) b.dof = (int)(unsigned char)(char)v & 63;
    else b.dof = (load(fr, "Fixed", v) && v) ? 63 : 0;  // older files: binary Fixed tag
    const char *fk[3] = {"ForceX", "ForceY", "ForceZ"}, *tk[3] = {"TorqueX", "TorqueY", "TorqueZ"};
    const char *dk[3] = {"DisplaceX", "DisplaceY", "DisplaceZ"};
    const char *ak[3] = {"AngDisplaceX", "AngDisplaceY", "AngDisplaceZ"};
    for (int k = 0; k < 3; k++) {
        load(fr, fk[k], b.force[k]);
        load(fr, tk[k], b.torque[k]);
        load(fr, dk[k], b.disp[k]);
        load(fr, ak[k], b.angDisp[k]);
    }
    if (b.prim == 3) {
        // CMesh::ReadXML (src/old/Mesh.cpp:133-217): vertices Vx, Vy, Vz and facets
        // V0, V1, V2 (absent values 0), then the bounding box
        const XNode *mesh = fr->child("CMesh");
        if (const XNode *vs = mesh ? mesh->child("Vertices") : nullptr)
            for (const XNode &vx : vs->kids) {
                if (vx.name != "Vertex") continue;
                double x = 0, y = 0, z = 0;
                load(&vx, "Vx", x); load(&vx, "Vy", y); load(&vx, "Vz", z);
                b.mv.insert(b.mv.end(), {x, y, z});
            }
        if (const XNode *fs = mesh ? mesh->child("Facets") : nullptr)
            for (const XNode &fc : fs->kids) {
                if (fc.name != "Facet") continue;
                int i0 = 0, i1 = 0, i2 = 0;
                load(&fc, "V0", i0); load(&fc, "V1", i1); load(&fc, "V2", i2);
                b.mf.insert(b.mf.end(), {i0, i1, i2});
            }
        const int nvert = (int)b.mv.size() / 3;
        for (int i : b.mf)
            if (i < 0 || i >= nvert) {  // the reference would read past its vertex list
                err = "mesh region facet with a vertex index out of range";
                return VX3MI_E_PARSE;
            }
        if (nvert) {  // CMesh::UpdateBoundingBox (Mesh.cpp:673-689)
            double lo[3] = {b.mv[0], b.mv[1], b.mv[2]}, hi[3] = {b.mv[0], b.mv[1], b.mv[2]};
            for (int i = 0; i < nvert; i++)
                for (int k = 0; k < 3; k++) {
                    const double x = b.mv[3 * i + k];
                    lo[k] = lo[k] < x ? lo[k] : x;
                    hi[k] = hi[k] > x ? hi[k] : x;
                }
            for (int k = 0; k < 3; k++) { b.bbMin[k] = lo[k]; b.bbSize[k] = hi[k] - lo[k]; }
        }
    }
    return VX3MI_OK;
}

// ---- mesh regions: the x-ray parity test of CMesh::IsInside -------------------
struct D3 { double x, y, z; };
D3 d3sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
double d3dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 d3cross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 d3norm(D3 a) {  // Vec3D::Normalized: unchanged when of zero length
    const double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 0 ? D3{a.x / l, a.y / l, a.z / l} : a;
}
double det3(D3 v0, D3 v1, D3 v2) {  // CMesh::Det (Mesh.cpp:877-884)
    return v0.x * (v1.y * v2.z - v1.z * v2.y) + v0.y * (v1.z * v2.x - v1.x * v2.z) + v0.z * (v1.x * v2.y - v1.y * v2.x);
}

// CMesh::IntersectXRay (Mesh.cpp:887-931): where the x-parallel ray at (y, z)
// meets the facet's plane, accepted by the two edge tests the reference makes
bool xray_hits(D3 a, D3 b, D3 c, double y, double z, D3 &p) {
    const double minX = a.x < b.x ? (a.x < c.x ? a.x : c.x) : (b.x < c.x ? b.x : c.x) - 1.0;
    const D3 o = {minX, y, z};
    const D3 n = d3norm(d3cross(d3sub(b, a), d3sub(c, a)));
    const double dn = 1.0 * n.x + 0.0 * n.y + 0.0 * n.z;
    if (std::fabs(dn) < 1E-5) return false;  // parallel
    const double dist = -d3dot(d3sub(o, a), n) / dn;
    p = D3{o.x + 1.0 * dist, o.y + 0.0 * dist, o.z + 0.0 * dist};
    const double V1 = d3dot(d3cross(d3sub(b, a), d3sub(p, a)), n);
    const double V2 = d3dot(d3cross(d3sub(c, b), d3sub(p, b)), n);
    return V1 >= 0 && V2 >= 0;
}

// CMesh::InsideTri (Mesh.cpp:851-874): p projects into the triangle
bool inside_tri(D3 p, D3 v0, D3 v1, D3 v2) {
    const D3 xax = d3norm(d3sub(v1, v0));
    const D3 zax = d3norm(d3cross(d3sub(v2, v0), xax));
    const D3 yax = d3norm(d3cross(zax, xax));
    const D3 p0 = {0, 0, 1};
    const D3 p1 = {d3dot(d3sub(v1, v0), xax), d3dot(d3sub(v1, v0), yax), 1};
    const D3 p2 = {d3dot(d3sub(v2, v0), xax), d3dot(d3sub(v2, v0), yax), 1};
    const D3 pt = {d3dot(d3sub(p, v0), xax), d3dot(d3sub(p, v0), yax), 1};
    const double e0 = det3(p0, p1, pt), e1 = det3(p1, p2, pt), e2 = det3(p2, p0, pt);
    return (e0 <= 0 && e1 <= 0 && e2 <= 0) || (e0 >= 0 && e1 >= 0 && e2 >= 0);
}

// CP_Mesh::IsIn (src/old/VX_FRegion.cpp:668-686) with CMesh::IsInside,
// FillTriLayer and FillTriLine (Mesh.cpp:787-804, 1046-1116): the point is
// mapped into mesh coordinates (region origin, uniform scale of the smallest
// region-to-mesh ratio, mesh bounding-box corner), and is inside when an odd
// number of the facets crossing its z plane and y line meet the x ray at or
// before it
bool mesh_is_in(const BcRegion &b, const double P[3], const double Env[3]) {
    if (b.mv.empty()) return false;
    double sc[3];
    for (int k = 0; k < 3; k++) sc[k] = (k == 0 ? b.dX : k == 1 ? b.dY : b.dZ) * (Env[k] / b.bbSize[k]);
    double s = sc[0];
    if (sc[1] < s) s = sc[1];
    if (sc[2] < s) s = sc[2];
    const double o[3] = {b.X * Env[0], b.Y * Env[1], b.Z * Env[2]};
    const D3 pt = {(P[0] - o[0]) / s + b.bbMin[0], (P[1] - o[1]) / s + b.bbMin[1], (P[2] - o[2]) / s + b.bbMin[2]};
    auto vert = [&](int i) { return D3{b.mv[3 * i], b.mv[3 * i + 1], b.mv[3 * i + 2]}; };
    std::vector<double> hits;
    for (size_t f = 0; f + 2 < b.mf.size(); f += 3) {
        const D3 a = vert(b.mf[f]), c1 = vert(b.mf[f + 1]), c2 = vert(b.mf[f + 2]);
        const bool below = a.z <= pt.z && c1.z <= pt.z && c2.z <= pt.z;  // FillTriLayer
        const bool above = a.z >= pt.z && c1.z >= pt.z && c2.z >= pt.z;
        if (below || above) continue;
        if (a.y < pt.y && c1.y < pt.y && c2.y < pt.y) continue;  // FillTriLine
        if (a.y > pt.y && c1.y > pt.y && c2.y > pt.y) continue;
        D3 p;
        if (xray_hits(a, c1, c2, pt.y, pt.z, p) && inside_tri(p, a, c1, c2)) hits.push_back(p.x);
    }
    std::sort(hits.begin(), hits.end());
    int count = 0;
    for (double x : hits) {
        if (pt.x < x) break;
        count++;
    }
    return count % 2 == 1;
}

// the first n FRegion children of a region list (CVX_Environment::ReadXML,
// src/old/VX_Environment.cpp:105-146)
int read_regions(const XNode *list, const char *countKey, std::vector<BcRegion> &out, std::string &err) {
    int n = 0;
    if (!load(list, countKey, n) || n <= 0) return VX3MI_OK;
    for (const XNode &k : list->kids) {
        if ((int)out.size() >= n) break;
        if (k.name != "FRegion") continue;
        BcRegion b;
        int rc = read_region(&k, b, err);
        if (rc != VX3MI_OK) return rc;
        out.push_back(b);
    }
    if ((int)out.size() < n) { err = "fewer FRegion elements than the region count"; return VX3MI_E_PARSE; }
    return VX3MI_OK;
}

// CPrimitive::IsTouching(P, Dist, Envelope) of the box / cylinder / sphere
// (src/old/VX_FRegion.cpp:420-428, 494-524, 601-605); a mesh is touched where
// its IsIn holds (:690-698)
bool region_touches(const BcRegion &b, const double P[3], const double D[3], const double Env[3]) {
    if (b.prim == 3) return mesh_is_in(b, P, Env);
    if (b.prim == 0) {
        const double B0[3] = {b.X, b.Y, b.Z}, B1[3] = {b.dX, b.dY, b.dZ};
        for (int k = 0; k < 3; k++) {
            const double ps = P[k] / Env[k], ds = D[k] / Env[k];
            if (!(ps + ds > B0[k] && ps - ds < B0[k] + B1[k])) return false;
        }
        return true;
    }
    const double pos[3] = {b.X * Env[0], b.Y * Env[1], b.Z * Env[2]};
    const double size[3] = {b.dX * Env[0], b.dY * Env[1], b.dZ * Env[2]};
    const double m01 = Env[0] > Env[1] ? Env[0] : Env[1], emax = Env[2] > m01 ? Env[2] : m01;  // Vec3D::Max()
    const double rad = b.radius * emax;
    if (b.prim == 2) {
        const double dist = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
        const double dx = P[0] - pos[0], dy = P[1] - pos[1], dz = P[2] - pos[2];
        return std::sqrt(dx * dx + dy * dy + dz * dz) < rad + dist;
    }
    // cylinder: IsIn along its one non-zero axis
    const double dim[3] = {b.dX, b.dY, b.dZ};
    for (int a = 0; a < 3; a++) {
        if (dim[a] == 0) continue;
        const int u = (a + 1) % 3, w = (a + 2) % 3;
        const int lo = u < w ? u : w, hi = u < w ? w : u;
        return P[a] > pos[a] && P[a] < pos[a] + size[a] &&
               std::sqrt((P[lo] - pos[lo]) * (P[lo] - pos[lo]) + (P[hi] - pos[hi]) * (P[hi] - pos[hi])) < rad;
    }
    return false;
}

// The boundary conditions of CVX_Sim::Import (src/VXA/VX_Sim.cpp:109-143): each
// region's voxel count (CVX_Environment::GetNumTouching, VX_Environment.cpp:215-232,
// from CVX_Object::GetXYZ's cell centres) divides its force and torque; a voxel
// whose centre touches the region gets the region's fixed degrees of freedom
// with their displacements, and force / moment added as
// external()->addForce(external()->force() + Force / count).
void apply_boundary_conditions(const VxaContent &c, Model &m) {
    const double Ld = c.latticeDim;
    const double env[3] = {Ld * c.dimAdj[0], Ld * c.dimAdj[1], Ld * c.dimAdj[2]};  // GetLatDimEnv()
    const double half[3] = {env[0] / 2.0, env[1] / 2.0, env[2] / 2.0};
    const double ws[3] = {Ld * c.dimAdj[0] * m.nx, Ld * c.dimAdj[1] * m.ny, Ld * c.dimAdj[2] * m.nz};
    const int nv = m.nv();
    std::vector<int> sizes(c.bcs.size(), 0);
    for (size_t j = 0; j < c.bcs.size(); j++)
        for (int v = 0; v < nv; v++) {
            const double eps = 0.000001;  // GetXYZ's offset (VX_Object.cpp:533)
            const double P[3] = {Ld * (c.dimAdj[0] * (0.5 + m.ix[v]) + eps), Ld * (c.dimAdj[1] * (0.5 + m.iy[v]) + eps),
                                 Ld * c.dimAdj[2] * (0.5 + m.iz[v])};
            if (region_touches(c.bcs[j], P, half, ws)) sizes[j]++;
        }
    const double s = env[0];  // voxel size (Vx.setVoxelSize(GetLatDimEnv().x))
    for (int v = 0; v < nv; v++) {
        const double P[3] = {m.ix[v] * s + half[0], m.iy[v] * s + half[1], m.iz[v] * s + half[2]};
        bool has = false;
        int dof = 0;
        double t[3] = {0, 0, 0}, r[3] = {0, 0, 0};
        float f[3] = {0, 0, 0}, mo[3] = {0, 0, 0};
        for (size_t j = 0; j < c.bcs.size(); j++) {
            const BcRegion &b = c.bcs[j];
            if (!region_touches(b, P, half, ws)) continue;
            has = true;
            for (int k = 0; k < 3; k++) {
                // setDisplacement: fixes the dof, keeps a zero displacement out
                if (b.dof & (1 << k)) { dof |= 1 << k; if (b.disp[k] != 0.0f) t[k] = b.disp[k]; }
                if (b.dof & (8 << k)) { dof |= 8 << k; if (b.angDisp[k] != 0.0f) r[k] = b.angDisp[k]; }
            }
            for (int k = 0; k < 3; k++) {
                const float gf = (float)((double)f[k] + b.force[k] / sizes[j]);
                const float gm = (float)((double)mo[k] + b.torque[k] / sizes[j]);
                f[k] += gf;
                mo[k] += gm;
            }
        }
        if (has) set_external(m, v, dof, t, r, f, mo);
    }
}

// boost property_tree get<bool> (VX3_SimulationManager.cu:328-330): "true"/"1" or "false"/"0"
bool load_bool(const XNode *parent, const char *key, bool dflt) {
    const XNode *n = parent ? parent->child(key) : nullptr;
    if (!n) return dflt;
    if (n->text == "true" || n->text == "1") return true;
    if (n->text == "false" || n->text == "0") return false;
    return dflt;
}

// VX3_SimulationManager::ParseMathTree (VX3_SimulationManager.cu:157-275):
// breadth-first walk below the formula element, then reversed, MT_END last
int parse_formula(const XNode *root, std::vector<MtToken> &out, std::string &err) {
    out.clear();
    if (!root || root->kids.empty()) return VX3MI_OK;
    static const struct { const char *name; int op; } ops[] = {
        {"mtCONST", MT_CONST}, {"mtE", MT_E}, {"mtPI", MT_PI}, {"mtVAR", MT_VAR}, {"mtADD", MT_ADD},
        {"mtSUB", MT_SUB}, {"mtMUL", MT_MUL}, {"mtDIV", MT_DIV}, {"mtPOW", MT_POW}, {"mtSQRT", MT_SQRT},
        {"mtSIN", MT_SIN}, {"mtCOS", MT_COS}, {"mtTAN", MT_TAN}, {"mtATAN", MT_ATAN}, {"mtLOG", MT_LOG},
        {"mtINT", MT_INT}, {"mtABS", MT_ABS}, {"mtNOT", MT_NOT}, {"mtGREATERTHAN", MT_GT},
        {"mtLESSTHAN", MT_LT}, {"mtAND", MT_AND}, {"mtOR", MT_OR}, {"mtNORMALCDF", MT_NORMALCDF}};
    static const char *vars[MV_COUNT] = {"x", "y", "z", "hit", "t", "angle", "targetCloseness", "numClosePairs",
                                         "num_voxel"};
    std::vector<const XNode *> bfs;
    for (const XNode &k : root->kids) bfs.push_back(&k);
    for (size_t head = 0; head < bfs.size(); head++)
        for (const XNode &k : bfs[head]->kids) bfs.push_back(&k);
    if (bfs.size() >= 1024) { err = "formula longer than 1023 tokens"; return VX3MI_E_PARSE; }
    for (auto it = bfs.rbegin(); it != bfs.rend(); ++it) {
        const XNode &n = **it;
        MtToken tk{-1, 0.0};
        for (const auto &o : ops)
            if (n.name == o.name) tk.op = o.op;
        if (tk.op < 0) { err = "unknown formula token <" + n.name + ">"; return VX3MI_E_PARSE; }
        if (tk.op == MT_VAR) {
            int var = -1;
            for (int v = 0; v < MV_COUNT; v++)
                if (n.text == vars[v]) var = v;
            if (var < 0) { err = "unknown formula variable '" + n.text + "'"; return VX3MI_E_PARSE; }
            tk.value = var;
        } else if (tk.op == MT_CONST) {
            char *end = nullptr;
            tk.value = std::strtod(n.text.c_str(), &end);
            if (end == n.text.c_str()) { err = "mtCONST without a number"; return VX3MI_E_PARSE; }
        }
        out.push_back(tk);
    }
    out.push_back(MtToken{MT_END, 0.0});
    const double ones[MV_COUNT] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
    double r;
    if (!mt_eval(out, ones, &r)) { err = "malformed formula (operator without operands)"; return VX3MI_E_PARSE; }
    return VX3MI_OK;
}

int read_vxa(const XNode &doc, VxaContent &c, std::string &err) {
    const XNode *vxa = doc.child("VXA");
    if (!vxa) { err = "no <VXA> root element"; return VX3MI_E_PARSE; }
    vx3mi_lattice &L = c.lat;
    std::memset(&L, 0, sizeof(L));
    if (const XNode *rp = vxa->child("RawPrint")) c.rawPrint = rp->text;  // VX3_SimulationManager.cu:322
    // ---- simulator (CVX_Sim::ReadXML; DtFrac defaults to 0.9 in CVX_Sim)
    L.dt_frac = 0.9;
    // damping: the reference leaves the zetas uninitialised without <Damping>; the
    // ReadXML defaults are used here
    L.bond_damp_z = 0.1f;
    L.col_damp_z = 1.0f;
    L.slow_damp_z = 1.0f;
    const XNode *sim = vxa->child("Simulator");
    if (sim) {
        if (const XNode *integ = sim->child("Integration")) {
            if (!load(integ, "DtFrac", L.dt_frac)) L.dt_frac = 0.9;
        }
        if (const XNode *damp = sim->child("Damping")) {
            if (!load(damp, "BondDampingZ", L.bond_damp_z)) L.bond_damp_z = 0.1f;
            if (!load(damp, "ColDampingZ", L.col_damp_z)) L.col_damp_z = 1.0f;
            if (!load(damp, "SlowDampingZ", L.slow_damp_z)) L.slow_damp_z = 1.0f;
        }
        if (const XNode *feat = sim->child("Features")) {
            int v = 0;
            if (load(feat, "VolumeEffectsEnabled", v)) L.volume_effects = v != 0;
        }
        // VX3 keys (VX3_SimulationManager.cu:323-376); EnableCollision defaults to true
        const XNode *ad = sim->child("AttachDetach");
        L.enable_collision = load_bool(ad, "EnableCollision", true);
        L.enable_attach = load_bool(ad, "EnableAttach", false);
        L.enable_detach = load_bool(ad, "EnableDetach", false);
        L.safety_guard = 500;
        load(ad, "SafetyGuard", L.safety_guard);
        int v;
        if (load(sim, "EnableExpansion", v)) L.enable_expansion = v;
        if (load(sim, "EnableCilia", v)) L.enable_cilia = v != 0;
        if (load(sim, "EnableSignals", v)) L.enable_signals = v != 0;
        double d;
        if (load(sim, "MaxDistInVoxelLengthsToCountAsPair", d)) L.max_dist_pairs = d;
        load(sim, "SecondaryExperiment", c.secondaryExperiment);
        L.secondary_experiment = c.secondaryExperiment != 0;
        load(sim, "SavePositionOfAllVoxels", c.savePositions);
        load(sim, "ReinitializeInitialPositionAfterThisManySeconds", c.reinitAfter);
        if (const XNode *rh = sim->child("RecordHistory")) {
            load(rh, "RecordStepSize", c.recordStepSize);
            load(rh, "RecordVoxel", c.recordVoxel);
            load(rh, "RecordLink", c.recordLink);
        }
        int rc = parse_formula(path(sim, {"StopCondition", "StopConditionFormula"}), c.stop, err);
        if (rc == VX3MI_OK) rc = parse_formula(sim->child("FitnessFunction"), c.fitness, err);
        if (rc != VX3MI_OK) return rc;
        // ForceField/{x,y,z}_forcefield (VX3_SimulationManager.cu:352-354); a
        // constant-zero tree is the reference's default and adds +0.0
        if (const XNode *ff = sim->child("ForceField")) {
            static const char *names[3] = {"x_forcefield", "y_forcefield", "z_forcefield"};
            for (int a = 0; a < 3; a++) {
                std::vector<MtToken> f, post;
                if ((rc = parse_formula(ff->child(names[a]), f, err)) != VX3MI_OK) return rc;
                if (f.empty() || (f.size() == 2 && f[0].op == MT_CONST && f[0].value == 0.0)) continue;
                if (!mt_postfix(f, post)) {
                    err = "force-field formula nested too deeply";
                    return VX3MI_E_PARSE;
                }
                c.forceField[a] = f;
            }
        }
        // AttachDetach/AttachCondition/Condition_0..4 (VX3_SimulationManager.cu:334-343)
        for (int k = 0; k < 5; k++) {
            const std::string name = "Condition_" + std::to_string(k);
            std::vector<MtToken> f, post;
            if ((rc = parse_formula(path(ad, {"AttachCondition", name.c_str()}), f, err)) != VX3MI_OK) return rc;
            if (f.empty()) continue;
            if (!mt_postfix(f, post)) {
                err = "attach condition nested too deeply";
                return VX3MI_E_PARSE;
            }
            c.attachCond[k] = f;
        }
    } else {
        L.enable_collision = 1;
        L.safety_guard = 500;
    }
    // ---- environment (CVX_Environment ctor defaults, src/old/VX_Environment.cpp:22-35)
    L.grav_acc = -9.81f;
    L.floor_enabled = 1;
    L.temp_enabled = 1;
    L.vary_temp = 0;
    L.temp_amplitude = 0;
    L.temp_base = 25;
    L.temp_period = 0.1;
    if (const XNode *env = vxa->child("Environment")) {
        int rc = VX3MI_OK;
        if (const XNode *bc = env->child("Boundary_Conditions")) {
            rc = read_regions(bc, "NumBCs", c.bcs, err);
        } else {  // older files: separate fixed and forced region lists
            if (const XNode *fx = env->child("Fixed_Regions")) rc = read_regions(fx, "NumFixed", c.bcs, err);
            if (rc == VX3MI_OK)
                if (const XNode *fo = env->child("Forced_Regions")) rc = read_regions(fo, "NumForced", c.bcs, err);
        }
        if (rc != VX3MI_OK) return rc;
        if (const XNode *g = env->child("Gravity")) {
            int v;
            double d;
            // GravEnabled is read but inert: SetGravityAccel always follows (VX_Sim.cpp:67-68, 485)
            L.grav_acc = load(g, "GravAcc", d) ? (float)d : -9.81f;
            L.floor_enabled = load(g, "FloorEnabled", v) ? (v != 0) : 0;
        }
        if (const XNode *t = env->child("Thermal")) {
            int v;
            double d;
            L.temp_enabled = load(t, "TempEnabled", v) ? (v != 0) : 0;
            L.temp_base = load(t, "TempBase", d) ? d : 25.0;
            if (load(t, "TempAmplitude", d)) L.temp_amplitude = d;
            else if (load(t, "TempAmp", d)) L.temp_amplitude = d - L.temp_base;
            else L.temp_amplitude = 0;
            L.vary_temp = load(t, "VaryTempEnabled", v) ? (v != 0) : 0;
            L.temp_period = load(t, "TempPeriod", d) ? d : 0.1;
        }
    }
    // ---- VXC
    const XNode *vxc = vxa->child("VXC");
    if (!vxc) vxc = vxa->child("DMF");
    if (!vxc) { err = "no <VXC> element"; return VX3MI_E_PARSE; }
    L.lattice_dim = 0.001;
    double xadj = 1.0;
    if (const XNode *lat = vxc->child("Lattice")) {
        if (!load(lat, "Lattice_Dim", L.lattice_dim)) L.lattice_dim = 0.001;
        if (!load(lat, "X_Dim_Adj", xadj)) xadj = 1.0;
        if (!load(lat, "Y_Dim_Adj", c.dimAdj[1])) c.dimAdj[1] = 1.0;
        if (!load(lat, "Z_Dim_Adj", c.dimAdj[2])) c.dimAdj[2] = 1.0;
        double off = 0;
        const char *offs[4] = {"X_Line_Offset", "Y_Line_Offset", "X_Layer_Offset", "Y_Layer_Offset"};
        for (const char *k : offs)
            if (!c.bcs.empty() && load(lat, k, off) && off != 0.0) {
                err = "boundary-condition regions on a lattice with line or layer offsets are not supported";
                return VX3MI_E_PARSE;
            }
    }
    c.latticeDim = L.lattice_dim;
    c.dimAdj[0] = xadj;
    L.lattice_dim *= xadj;  // Vx.setVoxelSize(LocalVXC.GetLatDimEnv().x) (VX_Sim.cpp:88)
    if (const XNode *pal = vxc->child("Palette")) {
        for (const XNode &mn : pal->kids) {
            if (mn.name != "Material") continue;
            vx3mi_material m;
            std::memset(&m, 0, sizeof(m));
            m.red = m.green = m.blue = 0.5f;
            m.alpha = 1.0f;
            m.is_measured = 1;
            m.signal_value_decay = 0.9;
            m.signal_time_delay = 0.03;
            m.inactive_period = 0.03;
            int matType = 0;
            if (load(&mn, "MatType", matType) && matType != 0) { err = "only MatType 0 (single) materials are supported"; return VX3MI_E_PARSE; }
            if (const XNode *disp = mn.child("Display")) {
                load(disp, "Red", m.red);
                load(disp, "Green", m.green);
                load(disp, "Blue", m.blue);
                load(disp, "Alpha", m.alpha);
            }
            if (const XNode *mech = mn.child("Mechanical")) {
                int model = 0, iv;
                load(mech, "MatModel", model);
                if (model < 0 || model > 3) { err = "MatModel must be 0, 1, 2 or 3"; return VX3MI_E_PARSE; }
                m.mat_model = model;
                // SSData (VX_Object.cpp:1397-1418): NumDataPts values under
                // StrainData/Strain and StressData/Stress
                if (const XNode *ss = mech->child("SSData")) {
                    int np = 0;
                    load(ss, "NumDataPts", np);
                    if (np > VX3MI_MAX_SS_POINTS) { err = "more than 32 stress-strain data points"; return VX3MI_E_LIMIT; }
                    std::vector<double> vals[2];
                    const char *grp[2] = {"StrainData", "StressData"}, *key[2] = {"Strain", "Stress"};
                    for (int g = 0; g < 2; g++)
                        if (const XNode *gd = ss->child(grp[g]))
                            for (const XNode &e : gd->kids)
                                if (e.name == key[g] && (int)vals[g].size() < np) vals[g].push_back(std::atof(e.text.c_str()));
                    if (np > 0 && ((int)vals[0].size() != np || (int)vals[1].size() != np)) {
                        err = "SSData holds fewer Strain or Stress values than NumDataPts";
                        return VX3MI_E_PARSE;
                    }
                    m.ss_points = np;
                    for (int i = 0; i < np; i++) {
                        m.ss_strain[i] = (float)vals[0][i];
                        m.ss_stress[i] = (float)vals[1][i];
                    }
                }
                if (model == 3 && m.ss_points < 1) {
                    // without data the reference reads past an empty vector (VX_Sim.cpp:400-407)
                    err = "MatModel 3 needs SSData points";
                    return VX3MI_E_PARSE;
                }
                load(mech, "Elastic_Mod", m.elastic_mod);
                load(mech, "Fail_Stress", m.fail_stress);
                load(mech, "Plastic_Mod", m.plastic_mod);
                load(mech, "Yield_Stress", m.yield_stress);
                load(mech, "Density", m.density);
                load(mech, "Poissons_Ratio", m.poissons_ratio);
                load(mech, "CTE", m.cte);
                load(mech, "uStatic", m.u_static);
                load(mech, "uDynamic", m.u_dynamic);
                if (load(mech, "isTarget", iv)) m.is_target = iv != 0;
                if (load(mech, "Fixed", iv)) m.fixed = iv != 0;
                if (load(mech, "Sticky", iv)) m.sticky = iv != 0;
                if (load(mech, "isMeasured", iv)) m.is_measured = iv;
                load(mech, "TurnOnThermalExpansionAfterThisManySeconds", m.thermal_on_after);
                load(mech, "Cilia", m.cilia);
                load(mech, "TurnOnCiliaAfterThisManySeconds", m.cilia_on_after);
                load(mech, "RemoveFromSimulationAfterThisManySeconds", m.remove_after);
                // signal keys and their ReadXML defaults (VX_Object.cpp:1420-1424)
                if (load(mech, "isPaceMaker", iv)) m.is_pacemaker = iv != 0;
                load(mech, "PaceMakerPeriod", m.pacemaker_period);
                load(mech, "signalValueDecay", m.signal_value_decay);
                load(mech, "signalTimeDelay", m.signal_time_delay);
                load(mech, "inactivePeriod", m.inactive_period);
            }
            c.mats.push_back(m);
        }
    }
    const XNode *st = vxc->child("Structure");
    if (!st) { err = "no <Structure> element"; return VX3MI_E_PARSE; }
    int nx = 1, ny = 1, nz = 1;
    load(st, "X_Voxels", nx);
    load(st, "Y_Voxels", ny);
    load(st, "Z_Voxels", nz);
    if (nx <= 0 || ny <= 0 || nz <= 0) { err = "bad structure dimensions"; return VX3MI_E_PARSE; }
    const char *comp = st->attr("Compression");
    std::string compression = comp ? comp : "";
    const XNode *data = st->child("Data");
    if (!data) { err = "no <Data> in <Structure>"; return VX3MI_E_PARSE; }
    c.cells.assign((size_t)nx * ny * nz, 0);
    int z = 0;
    for (const XNode &layer : data->kids) {
        if (layer.name != "Layer") continue;
        if (z >= nz) break;
        std::string raw = layer.text, in;
        if (compression == "ASCII_READABLE") {
            for (char ch : raw) in.push_back((char)(ch - 48));
        } else if (compression == "RAW_DATA") {
            in = raw;
        } else if (compression == "ZLIB" || compression == "QT_ZLIB") {
            err = "ZLIB-compressed layers are not supported (nor by the reference build)";
            return VX3MI_E_PARSE;
        } else {
            in = base64_decode(raw);
        }
        if ((long)in.size() != (long)nx * ny) { err = "voxel layer data does not match X_Voxels*Y_Voxels"; return VX3MI_E_PARSE; }
        for (int k = 0; k < nx * ny; k++) c.cells[(size_t)nx * ny * z + k] = (signed char)in[k];
        z++;
    }
    if (z < nz) { err = "fewer voxel layers than Z_Voxels"; return VX3MI_E_PARSE; }
    // PhaseOffset layers: one value per cell, kept for cells holding a voxel
    if (const XNode *po = st->child("PhaseOffset")) {
        int zz = 0;
        for (const XNode &layer : po->kids) {
            if (layer.name != "Layer") continue;
            if (zz >= nz) break;
            std::vector<std::string> vals = split_commas(layer.text);
            for (int k = 0; k < nx * ny; k++) {
                if (c.cells[(size_t)nx * ny * zz + k] > 0)
                    c.phase.push_back(k < (int)vals.size() ? std::atof(vals[k].c_str()) : 0.0);
            }
            zz++;
        }
    }
    // BaseCiliaForce / ShiftCiliaForce layers (VX_Object.cpp:1886-1958): three
    // values per cell, kept for cells holding a voxel; a layer too short for a
    // cell gives that cell zeros (as the reference pads)
    auto cilia_layers = [&](const char *name, std::vector<double> &out) {
        const XNode *cf = st->child(name);
        if (!cf) return;
        int zz = 0;
        for (const XNode &layer : cf->kids) {
            if (layer.name != "Layer") continue;
            if (zz >= nz) break;
            std::vector<std::string> vals = split_commas(layer.text);
            for (int k = 0; k < nx * ny; k++) {
                if (c.cells[(size_t)nx * ny * zz + k] <= 0) continue;
                for (int a = 0; a < 3; a++) {
                    const size_t i = (size_t)3 * k + a;
                    out.push_back(i < vals.size() ? std::atof(vals[i].c_str()) : 0.0);
                }
            }
            zz++;
        }
    };
    cilia_layers("BaseCiliaForce", c.ciliaBase);
    cilia_layers("ShiftCiliaForce", c.ciliaShift);
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    return VX3MI_OK;
}

int model_from_docs(const char *baseText, const char *vxdText, vx3mi_model **out) {
    if (!out) { set_error("out is NULL"); return VX3MI_E_ARG; }
    *out = nullptr;
    if (!baseText) { set_error("base VXA text is NULL"); return VX3MI_E_ARG; }
    VxaContent c;
    try {
        XNode doc = XmlReader(baseText).parse_document();
        if (vxdText && *vxdText) {
            XNode vxd = XmlReader(vxdText).parse_document();
            merge_vxd(doc, vxd);
        }
        std::string err;
        int rc = read_vxa(doc, c, err);
        if (rc != VX3MI_OK) { set_error(err); return rc; }
    } catch (const XmlError &e) {
        set_error("XML: " + e.msg);
        return VX3MI_E_PARSE;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return VX3MI_E_NOMEM;
    }
    c.lat.cells = c.cells.data();
    int nv = 0;
    for (signed char ch : c.cells) nv += (ch > 0 && ch <= (int)c.mats.size());
    c.lat.phase = (!c.phase.empty() && (int)c.phase.size() >= nv) ? c.phase.data() : nullptr;
    c.lat.voxel_E = nullptr;
    // missing layers leave zeros (the reference pads short data)
    if (!c.ciliaBase.empty()) c.ciliaBase.resize((size_t)3 * nv, 0.0);
    if (!c.ciliaShift.empty()) c.ciliaShift.resize((size_t)3 * nv, 0.0);
    c.lat.cilia_base = c.ciliaBase.empty() ? nullptr : c.ciliaBase.data();
    c.lat.cilia_shift = c.ciliaShift.empty() ? nullptr : c.ciliaShift.data();
    c.lat.nmat = (int)c.mats.size();
    c.lat.mats = c.mats.data();
    int rc = vx3mi_model_from_lattice(&c.lat, out);
    if (rc == VX3MI_OK) {
        Model &m = (*out)->m;
        m.rawPrint = c.rawPrint;
        m.savePositions = c.savePositions;
        m.stopFormula = c.stop;
        m.fitnessFormula = c.fitness;
        for (int a = 0; a < 3; a++) m.forceField[a] = c.forceField[a];
        for (int k = 0; k < 5; k++) m.attachCond[k] = c.attachCond[k];
        m.secondaryExperiment = c.secondaryExperiment;
        m.reinitInitialPositionAfter = c.reinitAfter;
        m.recordStepSize = c.recordStepSize;
        m.recordVoxel = c.recordVoxel;
        m.recordLink = c.recordLink;
        if (!c.bcs.empty()) apply_boundary_conditions(c, m);
    }
    return rc;
}

bool read_file(const char *path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

}  // namespace
}  // namespace vx3mi

using namespace vx3mi;

extern "C" {

int vx3mi_model_from_xml(const char *base_vxa_text, const char *vxd_text, vx3mi_model **out) {
    return model_from_docs(base_vxa_text, vxd_text, out);
}

int vx3mi_model_set_force_field(vx3mi_model *m, int axis, const char *formula_xml) {
    if (!m || axis < 0 || axis > 2) { set_error("vx3mi_model_set_force_field: bad model or axis"); return VX3MI_E_ARG; }
    std::vector<MtToken> f, post;
    try {
        if (formula_xml && *formula_xml) {
            // the formula element's content: the same subtree a VXA's <x_forcefield> holds
            const std::string wrapped = std::string("<f>") + formula_xml + "</f>";
            XNode doc = XmlReader(wrapped.c_str()).parse_document();
            std::string err;
            const int rc = parse_formula(doc.child("f"), f, err);
            if (rc != VX3MI_OK) { set_error(err); return rc; }
        }
    } catch (const XmlError &e) {
        set_error("XML: " + e.msg);
        return VX3MI_E_PARSE;
    }
    if (!f.empty() && !mt_postfix(f, post)) { set_error("force-field formula nested too deeply"); return VX3MI_E_PARSE; }
    m->m.forceField[axis] = f;
    return VX3MI_OK;
}

int vx3mi_model_set_attach_condition(vx3mi_model *m, int k, const char *formula_xml) {
    if (!m || k < 0 || k > 4) { set_error("vx3mi_model_set_attach_condition: bad model or condition index"); return VX3MI_E_ARG; }
    std::vector<MtToken> f, post;
    try {
        if (formula_xml && *formula_xml) {
            const std::string wrapped = std::string("<f>") + formula_xml + "</f>";
            XNode doc = XmlReader(wrapped.c_str()).parse_document();
            std::string err;
            const int rc = parse_formula(doc.child("f"), f, err);
            if (rc != VX3MI_OK) { set_error(err); return rc; }
        }
    } catch (const XmlError &e) {
        set_error("XML: " + e.msg);
        return VX3MI_E_PARSE;
    }
    if (!f.empty() && !mt_postfix(f, post)) { set_error("attach condition nested too deeply"); return VX3MI_E_PARSE; }
    m->m.attachCond[k] = f;
    return VX3MI_OK;
}

int vx3mi_model_from_vxa(const char *base_vxa_path, const char *vxd_path, vx3mi_model **out) {
    if (out) *out = nullptr;
    if (!base_vxa_path) { set_error("base VXA path is NULL"); return VX3MI_E_ARG; }
    std::string base, vxd;
    if (!read_file(base_vxa_path, base)) { set_error(std::string("cannot read ") + base_vxa_path); return VX3MI_E_PARSE; }
    if (vxd_path && !read_file(vxd_path, vxd)) { set_error(std::string("cannot read ") + vxd_path); return VX3MI_E_PARSE; }
    return model_from_docs(base.c_str(), vxd_path ? vxd.c_str() : nullptr, out);
}

}  // extern "C"
