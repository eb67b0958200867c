// metriko (MI355X build) -- mesh and parameterization files.
//
// read_obj: the `v x y z` and triangle `f a b c` lines of a Wavefront OBJ (1-based indices,
// `a/b/c` and `a//c` forms accepted), as igl::readOBJ fills V (nV x 3) and F (nF x 3) for
// the reference's example meshes (example_igm.cpp:16, example_qex.cpp:50).
// write_cfn_json / read_cfn_json: the corner-uv dump of example_qex.cpp:14-43 (toJson /
// fromJson, commented out there): one JSON object {"cfn": [re0, im0, re1, im1, ...]}; values
// are written with 17 significant digits, so a round trip is exact.
#ifndef METRIKO_COMMON_IO_H
#define METRIKO_COMMON_IO_H
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#include "typedef.h"

namespace metriko {
    inline bool read_obj(const std::string& path, MatXd& V, MatXi& F) {
        std::ifstream in(path);
        if (!in) return false;
        std::vector<double> vs;
        std::vector<int> fs;
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            std::string tag;
            ss >> tag;
            if (tag == "v") {
                double x, y, z;
                if (!(ss >> x >> y >> z)) throw std::runtime_error("read_obj: bad vertex line: " + line);
                vs.insert(vs.end(), {x, y, z});
            } else if (tag == "f") {
                std::string t;
                int k = 0;
                while (ss >> t) {
                    if (k == 3) throw std::runtime_error("read_obj: only triangles are read: " + line);
                    fs.push_back(std::atoi(t.substr(0, t.find('/')).c_str()) - 1);
                    k++;
                }
                if (k != 3) throw std::runtime_error("read_obj: bad face line: " + line);
            }
        }
        const int nV = int(vs.size() / 3), nF = int(fs.size() / 3);
        V.resize(nV, 3);
        F.resize(nF, 3);
        for (size_t i = 0; i < vs.size(); i++) V.d[i] = vs[i];
        for (size_t i = 0; i < fs.size(); i++) {
            if (fs[i] < 0 || fs[i] >= nV) throw std::runtime_error("read_obj: face index out of range");
            F.d[i] = fs[i];
        }
        return true;
    }

    inline void write_cfn_json(const std::string& path, const VecXc& cfn) {
        std::FILE* f = std::fopen(path.c_str(), "w");
        if (!f) throw std::runtime_error("write_cfn_json: cannot open " + path);
        std::fputs("{\"cfn\":[", f);
        for (int i = 0; i < cfn.size(); i++)
            std::fprintf(f, "%s%.17g,%.17g", i ? "," : "", cfn[i].real(), cfn[i].imag());
        std::fputs("]}\n", f);
        std::fclose(f);
    }

    inline bool read_cfn_json(const std::string& path, MatXd& uv1, VecXc& uv2) {
        std::ifstream in(path);
        if (!in) return false;
        const std::string s((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        const size_t key = s.find("\"cfn\"");
        if (key == std::string::npos) throw std::runtime_error("read_cfn_json: no \"cfn\" array");
        size_t p = s.find('[', key);
        const size_t end = s.find(']', p);
        if (p == std::string::npos || end == std::string::npos) throw std::runtime_error("read_cfn_json: bad array");
        std::vector<double> data;
        for (p++; p < end;) {
            while (p < end && (s[p] == ',' || std::isspace(static_cast<unsigned char>(s[p])))) p++;
            if (p >= end) break;
            char* q = nullptr;
            data.push_back(std::strtod(s.c_str() + p, &q));
            if (q == s.c_str() + p) throw std::runtime_error("read_cfn_json: bad number");
            p = size_t(q - s.c_str());
        }
        if (data.size() % 2) throw std::runtime_error("read_cfn_json: odd number of values");
        const int n = int(data.size() / 2);
        uv1.resize(n, 2);
        uv2.resize(n);
        for (int i = 0; i < n; i++) {
            uv1(i, 0) = data[2 * size_t(i)];
            uv1(i, 1) = data[2 * size_t(i) + 1];
            uv2(i) = complex(data[2 * size_t(i)], data[2 * size_t(i) + 1]);
        }
        return true;
    }
}
#endif
