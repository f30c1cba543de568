// rad_obj.cpp - Wavefront OBJ faces as source quads (host code).
//
// Restates WaveFrontModel::parse (WaveFrontModel.cpp:15-136): the file is
// read line by line; only lines starting with "v " and "f " are used.
//  - "v x y z [w]": the numbers are read as floats until one fails; fewer
//    than three ends the parse (the faces read so far are kept, the call
//    reports RAD_E_ARG); a fourth is ignored.  The vertex is (x, y, z) / 1000
//    (the reference's millimetre models).
//  - "f a b c [d]": the line is cut at single spaces; each piece's leading
//    integer (anything after it, such as "/2/3", is ignored; a piece with
//    none is skipped) minus one is a vertex index.  At most five pieces are
//    taken.  Three vertices become the quad A B C C; a face that then does
//    not have four vertices, or names a vertex not yet defined (0, negative
//    or too large), is skipped.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/radhip.h"
#include "../../include/radhip_scene.h"

extern "C" {

int rad_obj_parse(const char *path, float **quads, uint32_t *nquad)
{
    if (!path || !quads || !nquad)
        return RAD_E_ARG;
    *quads = nullptr;
    *nquad = 0;
    std::ifstream f(path);
    if (!f.is_open())
        return RAD_E_ARG;
    std::vector<float> verts, out;
    std::string line;
    int rc = RAD_OK;
    while (f.good()) {
        std::getline(f, line);
        if (line.compare(0, 2, "v ") == 0) {
            std::istringstream in(line.substr(2));
            std::vector<float> p;
            float x;
            while (in >> x)
                p.push_back(x);
            if (p.size() < 3) {
                rc = RAD_E_ARG;
                break;
            }
            for (int k = 0; k < 3; ++k)
                verts.push_back(p[k] / 1000);
            continue;
        }
        if (line.compare(0, 2, "f ") == 0) {
            std::string rest = line.substr(2);
            std::vector<uint32_t> fv;
            while (!rest.empty() && fv.size() <= 4) {
                size_t pos = rest.find(' ');
                std::istringstream in(rest.substr(0, pos));
                int idx;
                if (in >> idx)
                    fv.push_back((uint32_t)(idx - 1));
                rest = pos == std::string::npos ? std::string() : rest.substr(pos + 1);
            }
            if (fv.size() == 3)
                fv.push_back(fv.back());
            if (fv.size() != 4)
                continue;
            size_t nv = verts.size() / 3;
            if (fv[0] >= nv || fv[1] >= nv || fv[2] >= nv || fv[3] >= nv)
                continue;
            for (uint32_t i : fv)
                out.insert(out.end(), verts.begin() + 3 * (size_t)i, verts.begin() + 3 * (size_t)i + 3);
        }
    }
    size_t n = out.size() / 12;
    if (n > 0xFFFFFFFFull)
        return RAD_E_LIMIT;
    float *q = (float *)malloc((out.size() + 1) * sizeof(float));
    if (!q)
        return RAD_E_NOMEM;
    if (!out.empty())
        memcpy(q, out.data(), out.size() * sizeof(float));
    *quads = q;
    *nquad = (uint32_t)n;
    return rc;
}

void rad_obj_free(float *quads) { free(quads); }

} // extern "C"
