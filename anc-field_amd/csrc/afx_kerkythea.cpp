// Kerkythea scene loader: axis-aligned boxes -> voxel grid + material filters.
//
// Restates anc-field/src/kerkythea.hpp:63-196 (LoadBoxes and the Kerkythea
// constructor) and anc-field/src/util.hpp:43-47 (RoundTo), :101-112
// (FileToVariable), with a small XML reader of its own in place of pugixml.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/afx.h"

namespace afx {

namespace {

struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XNode>> kids;
    const char* attr(const char* k) const
    {
        for (auto& a : attrs)
            if (a.first == k) return a.second.c_str();
        return "";
    }
    const XNode* child(const char* n) const
    {
        for (auto& c : kids)
            if (c->name == n) return c.get();
        return nullptr;
    }
};

std::string unescape(const std::string& s)
{
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '&') {
            const size_t e = s.find(';', i);
            if (e != std::string::npos) {
                const std::string ent = s.substr(i + 1, e - i - 1);
                if (ent == "lt") o += '<';
                else if (ent == "gt") o += '>';
                else if (ent == "amp") o += '&';
                else if (ent == "quot") o += '"';
                else if (ent == "apos") o += '\'';
                else o += s.substr(i, e - i + 1);
                i = e;
                continue;
            }
        }
        o += s[i];
    }
    return o;
}

// Element tree of an XML document (tags and attributes only; text ignored).
bool parse_xml(const std::string& t, XNode& root, std::string& err)
{
    std::vector<XNode*> stack{&root};
    size_t i = 0;
    while (true) {
        i = t.find('<', i);
        if (i == std::string::npos) break;
        if (t.compare(i, 4, "<!--") == 0) {
            i = t.find("-->", i);
            if (i == std::string::npos) { err = "unterminated comment"; return false; }
            i += 3;
            continue;
        }
        if (t.compare(i, 2, "<?") == 0 || t.compare(i, 2, "<!") == 0) {
            i = t.find('>', i);
            if (i == std::string::npos) break;
            ++i;
            continue;
        }
        if (t.compare(i, 2, "</") == 0) {
            const size_t e = t.find('>', i);
            if (e == std::string::npos) { err = "unterminated end tag"; return false; }
            if (stack.size() > 1) stack.pop_back();
            i = e + 1;
            continue;
        }
        size_t j = i + 1;
        while (j < t.size() && !isspace((unsigned char)t[j]) && t[j] != '>' && t[j] != '/') ++j;
        auto node = std::make_unique<XNode>();
        node->name = t.substr(i + 1, j - i - 1);
        bool selfclose = false;
        while (j < t.size()) {
            while (j < t.size() && isspace((unsigned char)t[j])) ++j;
            if (j >= t.size()) { err = "unterminated tag"; return false; }
            if (t[j] == '>') { ++j; break; }
            if (t[j] == '/') { selfclose = true; j = t.find('>', j); if (j == std::string::npos) { err = "bad tag"; return false; } ++j; break; }
            size_t k = j;
            while (k < t.size() && t[k] != '=' && !isspace((unsigned char)t[k]) && t[k] != '>') ++k;
            std::string key = t.substr(j, k - j);
            while (k < t.size() && isspace((unsigned char)t[k])) ++k;
            if (k >= t.size() || t[k] != '=') { j = k; continue; }
            ++k;
            while (k < t.size() && isspace((unsigned char)t[k])) ++k;
            if (k >= t.size() || (t[k] != '"' && t[k] != '\'')) { err = "unquoted attribute"; return false; }
            const char q = t[k];
            const size_t e = t.find(q, k + 1);
            if (e == std::string::npos) { err = "unterminated attribute"; return false; }
            node->attrs.push_back({key, unescape(t.substr(k + 1, e - k - 1))});
            j = e + 1;
        }
        XNode* raw = node.get();
        stack.back()->kids.push_back(std::move(node));
        if (!selfclose) stack.push_back(raw);
        i = j;
    }
    return true;
}

// util.hpp:101-112: one value per line, std::stod, cast to float
bool file_to_values(const std::string& path, std::vector<float>& out)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    out.clear();
    while (std::getline(f, line)) {
        size_t used = 0;
        try {
            out.push_back((float)std::stod(line, &used));
        } catch (...) {
            if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
            return false;
        }
    }
    return true;
}

uint32_t round_to(uint32_t n, uint32_t unit) { return n % unit == 0 ? n : n + (unit - n % unit); }

}  // namespace

int load_kerkythea(const char* path, float res, int32_t dims[3], std::vector<int32_t>& vox, std::vector<float>& a,
                   std::vector<float>& b, int32_t* ncoef, int32_t* nmat, std::string& err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = std::string("cannot open ") + path;
        return AFX_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    XNode doc;
    std::string perr;
    if (!parse_xml(ss.str(), doc, perr)) {
        err = std::string(path) + ": " + perr;
        return AFX_E_IO;
    }
    struct Box {
        float v[6];
        std::string mat;
    };
    std::vector<Box> boxes;
    // kerkythea.hpp:70-123
    const XNode* rootn = doc.child("Root");
    if (rootn)
        for (auto& scene : rootn->kids) {
            if (scene->name != "Object" || std::string(scene->attr("Type")) != "Scene") continue;
            for (auto& model : scene->kids) {
                if (model->name != "Object" || std::string(model->attr("Type")) != "Model") continue;
                const std::string name = model->attr("Name");
                const XNode* mesh = model->child("Object");
                if (!mesh) continue;
                for (auto& par : mesh->kids) {
                    if (par->name != "Parameter" || std::string(par->attr("Name")) != "Vertex List") continue;
                    float mnx = 10000.0f, mny = 10000.0f, mnz = 10000.0f;
                    float mxx = -10000.0f, mxy = -10000.0f, mxz = -10000.0f;
                    for (auto& p : par->kids) {
                        if (p->name != "P") continue;
                        float pa = 0, pb = 0, pc = 0;
                        std::sscanf(p->attr("xyz"), "%f %f %f", &pa, &pb, &pc);
                        const float x = pa, y = pc, z = -pb;  // axis remap, kerkythea.hpp:97-99
                        if (x < mnx) mnx = x;
                        if (x > mxx) mxx = x;
                        if (y < mny) mny = y;
                        if (y > mxy) mxy = y;
                        if (z < mnz) mnz = z;
                        if (z > mxz) mxz = z;
                    }
                    const std::string mat = name.substr(name.find_last_of('_') + 1);
                    if (mnx < 0.0f) mnx = 0.0f;
                    if (mny < 0.0f) mny = 0.0f;
                    if (mnz < 0.0f) mnz = 0.0f;
                    boxes.push_back({{mnx, mxx, mny, mxy, mnz, mxz}, mat});
                }
            }
        }
    if (boxes.empty()) {
        err = std::string(path) + ": no boxes found";
        return AFX_E_IO;
    }
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    for (auto& bx : boxes) {
        if (bx.v[1] > mx) mx = bx.v[1];
        if (bx.v[3] > my) my = bx.v[3];
        if (bx.v[5] > mz) mz = bx.v[5];
    }
    const uint32_t sx = round_to((uint32_t)(size_t)(mx / res) + 1, 32);
    const uint32_t sy = round_to((uint32_t)(size_t)(my / res) + 1, 32);
    const uint32_t sz = round_to((uint32_t)(size_t)(mz / res) + 1, 32);
    // materials in lexicographic order (std::set), kerkythea.hpp:157-177
    std::set<std::string> names;
    for (auto& bx : boxes) names.insert(bx.mat);
    std::string dir = path;
    const size_t slash = dir.find_last_of("/\\");
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    std::map<std::string, int32_t> ids;
    a.clear();
    b.clear();
    int32_t n = -1, k = 0;
    for (auto& nm : names) {
        std::vector<float> va, vb;
        if (!file_to_values(dir + "/" + nm + "_a.dat", va) || !file_to_values(dir + "/" + nm + "_b.dat", vb)) {
            err = "material " + nm + ": cannot read " + dir + "/" + nm + "_a.dat / _b.dat";
            return AFX_E_IO;
        }
        if (n < 0) n = (int32_t)va.size();
        if ((int32_t)va.size() != n || (int32_t)vb.size() != n || n == 0) {
            err = "material " + nm + ": coefficient counts differ from the first material's " + std::to_string(n);
            return AFX_E_IO;
        }
        a.insert(a.end(), va.begin(), va.end());
        b.insert(b.end(), vb.begin(), vb.end());
        ids[nm] = k++;
    }
    *ncoef = n;
    *nmat = k;
    dims[0] = (int32_t)sx;
    dims[1] = (int32_t)sy;
    dims[2] = (int32_t)sz;
    vox.assign((size_t)sx * sy * sz, -1);
    // inclusive index ranges, kerkythea.hpp:181-195
    for (auto& bx : boxes) {
        const int32_t id = ids[bx.mat];
        const size_t z0 = (size_t)(bx.v[4] / res), z1 = (size_t)(bx.v[5] / res);
        const size_t y0 = (size_t)(bx.v[2] / res), y1 = (size_t)(bx.v[3] / res);
        const size_t x0 = (size_t)(bx.v[0] / res), x1 = (size_t)(bx.v[1] / res);
        for (size_t z = z0; z <= z1 && z < sz; ++z)
            for (size_t y = y0; y <= y1 && y < sy; ++y)
                for (size_t x = x0; x <= x1 && x < sx; ++x) vox[(z * sy + y) * sx + x] = id;
    }
    return AFX_OK;
}

}  // namespace afx
