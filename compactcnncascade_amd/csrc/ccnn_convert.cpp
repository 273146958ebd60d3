// Conversion of a CNTK model dump (the text that CNTK's dumpNode writes for a trained
// cascade network) into the binary model format v1.1 that the detector loads
// (FaceDetector::CNTKDump2Binary, CompactCNNLibAPI_v2.cpp:509-519; the reference's
// CNNModelsConverter, cnn_models_converter.cpp:79-746).
//
// The dump is a list of nodes. A parameter node is a header line "z.<node>.<param>=..."
// followed by lines of numbers up to a line containing '#'; "._" is dropped from every
// line first. The network's shape comes from which node names exist (z.c<l><m> for map m
// of conv layer l, z.h<g> for hidden-layer group g) and from the "Kernel:" fields of the
// convolution headers. The conversion folds batch normalisation into the LReLU weights,
// sums the three input channels of the first-layer kernels and folds the input
// normalisation (x - 128) / 128 into them, and transposes the kernels to row-major order.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ccnn.h"

namespace {

struct Dump {
    std::vector<std::string> lines;

    bool load(const char* path)
    {
        std::ifstream f(path);
        if (!f.is_open()) return false;
        std::string s;
        while (std::getline(f, s)) {
            for (size_t i = s.find("._"); i != std::string::npos; i = s.find("._")) s.erase(i, 2);
            lines.push_back(s);
        }
        return true;
    }

    int find(const std::string& key) const
    {
        for (size_t i = 0; i < lines.size(); ++i)
            if (lines[i].find(key) != std::string::npos) return (int)i;
        return -1;
    }
    bool has(const std::string& key) const { return find(key) >= 0; }

    static std::vector<std::string> tokens(const std::string& s)
    {
        std::istringstream in(s);
        std::vector<std::string> t;
        std::string w;
        while (in >> w) t.push_back(w);
        return t;
    }

    // integer fields of the first line containing `key`, counted from `anchor` in it
    bool fields(const std::string& key, const std::string& anchor, int i0, int i1, int* a, int* b) const
    {
        const int l = find(key);
        if (l < 0) return false;
        const size_t p = lines[l].find(anchor);
        if (p == std::string::npos) return false;
        const std::vector<std::string> t = tokens(lines[l].substr(p));
        if ((int)t.size() <= std::max(i0, i1)) return false;
        *a = std::atoi(t[i0].c_str());
        *b = std::atoi(t[i1].c_str());
        return true;
    }

    // number of n = 1, 2, ... for which prefix + n appears somewhere in the dump
    int count(const std::string& prefix) const
    {
        int n = 1;
        while (has(prefix + std::to_string(n))) ++n;
        return n - 1;
    }

    // the numbers of parameter node z.<node>.<param>: every line after its header up to
    // a line containing '#', or only line `row` of them
    bool values(const std::string& node, const std::string& param, std::vector<float>& out, int row = -1) const
    {
        out.clear();
        const int l = find("z." + node + "." + param + "=");
        if (l < 0) return false;
        for (size_t i = l + 1, k = 0; i < lines.size() && lines[i].find('#') == std::string::npos; ++i, ++k) {
            if (row >= 0 && (int)k != row) continue;
            for (const std::string& w : tokens(lines[i])) out.push_back(std::strtof(w.c_str(), nullptr));
        }
        return !out.empty();
    }
};

struct Model {
    int min_w = 0, min_h = 0;
    bool max_pool = false;
    std::vector<int> maps;
    int kw[3] = {0, 0, 0}, kh[3] = {0, 0, 0};
    std::vector<std::vector<float>> kern[3];
    float af_scale = 1.7159f;
    std::vector<float> conv_b[3], w1[3], w2[3], bn_w[3], bn_b[3];
    bool full_connect = false;
    int hl = 0, connect = 0, hl_scale = 0, hl_nw = 0;
    std::vector<std::vector<float>> hl_w;
    std::vector<float> hl_b, hl_tanh_w, hl_bn_w, hl_bn_b;
    std::vector<std::vector<float>> ol_w;
    std::vector<float> ol_b;
    float ol_tanh_w = 0.f;
};

// kernel read from the dump's column order into row-major order
std::vector<float> transpose(const std::vector<float>& p, int kw, int kh)
{
    std::vector<float> t((size_t)kw * kh);
    for (int n = 0; n < kw * kh; ++n) t[n] = p[(size_t)kh * (n % kw) + n / kw];
    return t;
}

bool convert(const Dump& d, Model& m)
{
    // input size and pooling (first "features" line: "[ w x h ...", first "Pooling" line:
    // the pooling kind is the second field of the next line; 1 is max)
    if (!d.fields("features", "[", 1, 3, &m.min_w, &m.min_h)) return false;
    const int pl = d.find("Pooling");
    if (pl < 0 || pl + 1 >= (int)d.lines.size()) return false;
    const std::vector<std::string> pt = Dump::tokens(d.lines[pl + 1]);
    if (pt.size() < 2) return false;
    m.max_pool = std::atoi(pt[1].c_str()) == 1;

    const int layers = d.count("z.c");
    if (layers != 3) return false;                     // the detector's networks have three
    for (int l = 1; l <= layers; ++l) m.maps.push_back(d.count("z.c" + std::to_string(l)));
    std::vector<float> p;

    // layer 1: the three colour channels summed, normalisation (x - 128) / 128 folded in
    std::vector<double> shift(m.maps[0], 0.0);
    for (int l = 0; l < 3; ++l) {
        if (!d.fields("z.c" + std::to_string(l + 1) + "1.c=", "Kernel:", 1, 3, &m.kw[l], &m.kh[l])) return false;
        const int kw = m.kw[l], kh = m.kh[l];
        if (kw <= 0 || kh <= 0) return false;
        for (int k = 0; k < m.maps[l]; ++k) {
            if (!d.values("c" + std::to_string(l + 1) + std::to_string(k + 1), "W", p)) return false;
            std::vector<float> t;
            if (l == 0) {
                if (p.size() < (size_t)3 * kw * kh) return false;
                std::vector<float> sum((size_t)kw * kh);
                for (int y = 0; y < kh; ++y)
                    for (int x = 0; x < kw; ++x)
                        sum[(size_t)y * kw + x] = p[(size_t)3 * kw * y + x] + p[(size_t)3 * kw * y + kw + x] +
                                                  p[(size_t)3 * kw * y + 2 * kw + x];
                t = transpose(sum, kw, kh);
                for (size_t i = 0; i < t.size(); ++i) {
                    shift[k] += -1.0 * (double)t[i];
                    t[i] *= 0.0078125f;
                }
            } else {
                if (p.size() < (size_t)kw * kh) return false;
                t = transpose(p, kw, kh);
            }
            m.kern[l].push_back(t);
        }
    }

    // per map: conv bias, LReLU weights, batch normalisation folded into them
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < m.maps[l]; ++k) {
            const std::string c = "c" + std::to_string(l + 1) + std::to_string(k + 1);
            const std::string bn = "bn" + std::to_string(l + 1) + std::to_string(k + 1);
            float v[7];
            const char* names[7][2] = {{"", "res.PlusArgs[0].ElementTimesArgs[0]"},
                                       {"", "res.PlusArgs[1].ElementTimesArgs[0]"},
                                       {"", "b"},
                                       {"bn", "scale"}, {"bn", "bias"}, {"bn", "runMean"}, {"bn", "runVariance"}};
            for (int i = 0; i < 7; ++i) {
                if (!d.values(names[i][0][0] ? bn : c, names[i][1], p)) return false;
                v[i] = p[0];
            }
            const float var = v[6] + 0.00001f;
            const float bw = v[3] / sqrtf(var);
            const float bb = v[4] - v[5] * v[3] / sqrtf(var);
            m.w1[l].push_back(v[0] * bw);
            m.w2[l].push_back(v[1] * bw);
            m.bn_w[l].push_back(bw);
            m.bn_b[l].push_back(bb);
            m.conv_b[l].push_back(l == 0 ? v[2] + (float)shift[k] : v[2]);
        }

    // hidden layer: z.h<g> groups; "Kernel: connect x . x . x scale" of z.h1's convolution
    const int groups = d.count("z.h");
    int f1, f3;
    if (groups <= 0 || !d.fields("z.h1.c=Convolution", "Kernel:", 1, 7, &m.connect, &m.hl_scale) ||
        !d.fields("z.h1.c=Convolution", "Kernel:", 1, 3, &f1, &f3) || m.connect <= 0 || m.hl_scale <= 0)
        return false;
    m.hl = groups * m.hl_scale;
    m.full_connect = f1 == m.maps[2] || f3 == m.maps[2];
    // a fully connected hidden layer is written with one weight per last-layer map
    m.hl_nw = m.full_connect ? m.maps[2] : m.connect;
    if (m.hl_nw > m.connect) return false;
    m.hl_w.assign(m.hl, std::vector<float>(m.connect));
    m.hl_b.assign(m.hl, 0.f);
    for (int g = 0; g < groups; ++g) {
        const std::string h = "h" + std::to_string(g + 1);
        if (!d.values(h, "W", p) || p.size() < (size_t)m.connect * m.hl_scale) return false;
        for (int j = 0; j < m.hl_scale; ++j)
            for (int k = 0; k < m.connect; ++k) {
                const int n = j * m.connect + k;
                m.hl_w[(size_t)m.hl_scale * g + j][k] = p[(size_t)m.hl_scale * (n % m.connect) + n / m.connect];
            }
        if (!d.values(h, "b", p) || (int)p.size() < m.hl_scale) return false;
        for (int j = 0; j < m.hl_scale; ++j) m.hl_b[(size_t)m.hl_scale * g + j] = p[j];
        if (!d.values(h, "res.PlusArgs[0].ElementTimesArgs[1].x.ElementTimesArgs[0]", p)) return false;
        m.hl_tanh_w.push_back(p[0]);
    }
    for (int j = 0; j < m.hl_scale; ++j) {
        float v[4];
        const char* names[4] = {"scale", "bias", "runMean", "runVariance"};
        for (int i = 0; i < 4; ++i) {
            if (!d.values("bn_h", names[i], p, j)) return false;
            v[i] = p[0];
        }
        const float var = v[3] + 0.00001f;
        m.hl_bn_w.push_back(v[0] / sqrtf(var));
        m.hl_bn_b.push_back(v[1] - v[2] * v[0] / sqrtf(var));
    }

    // output layer: W is [outputs][hl]; more than three outputs are read as probabilities
    // (af_scale 0: 0.5 * tanh + 0.5)
    if (!d.values("z", "arrayOfFunctions[0].W", p) || p.size() < (size_t)m.hl) return false;
    const size_t outs = p.size() / m.hl;
    for (size_t i = 0; i < outs; ++i) m.ol_w.emplace_back(p.begin() + i * m.hl, p.begin() + (i + 1) * m.hl);
    if (!d.values("z", "arrayOfFunctions[0].b", p)) return false;
    m.ol_b = p;
    if (m.ol_b.size() != outs) return false;
    if (outs > 3) m.af_scale = 0.f;
    if (!d.values("z", "PlusArgs[0].ElementTimesArgs[1].x.ElementTimesArgs[0]", p)) return false;
    m.ol_tanh_w = p[0];
    return true;
}

template <class T>
void put(std::ofstream& f, T v)
{
    f.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

// binary format v1.1 (cnn_simd_cntk.cpp:71-394)
bool save(const Model& m, std::string path)
{
    if (path.find(".bin") == std::string::npos) path += ".bin";
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    put<float>(f, 1.1f);
    put<unsigned char>(f, m.max_pool ? 1 : 0);
    put<int>(f, m.min_w);
    put<int>(f, m.min_h);
    put<int>(f, (int)m.maps.size());
    for (int c : m.maps) put<int>(f, c);
    for (int l = 0; l < 3; ++l) {
        put<int>(f, m.kw[l]);
        put<int>(f, m.kh[l]);
        for (const auto& k : m.kern[l])
            for (float v : k) put<float>(f, v);
    }
    put<float>(f, std::fabs(m.af_scale));
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < m.maps[l]; ++k) {
            put<float>(f, m.conv_b[l][k]);
            put<float>(f, m.w1[l][k]);
            put<float>(f, m.w2[l][k]);
            put<float>(f, m.bn_w[l][k]);
            put<float>(f, m.bn_b[l][k]);
        }
    put<unsigned char>(f, m.full_connect ? 1 : 0);
    put<int>(f, m.hl);
    put<int>(f, m.connect);
    put<int>(f, m.hl_scale);
    for (const auto& w : m.hl_w)
        for (int k = 0; k < m.hl_nw; ++k) put<float>(f, w[k]);
    for (float v : m.hl_b) put<float>(f, v);
    for (float v : m.hl_tanh_w) put<float>(f, v);
    for (int j = 0; j < m.hl_scale; ++j) {
        put<float>(f, m.hl_bn_w[j]);
        put<float>(f, m.hl_bn_b[j]);
    }
    put<int>(f, (int)m.ol_w.size());
    for (const auto& w : m.ol_w)
        for (float v : w) put<float>(f, v);
    for (float v : m.ol_b) put<float>(f, v);
    put<float>(f, m.ol_tanh_w);
    return (bool)f;
}

}  // namespace

extern "C" int ccnn_cntk_dump_to_binary(const char* binary_file, const char* cntk_model_dump)
{
    if (!binary_file || !cntk_model_dump) return -1;
    Dump d;
    if (!d.load(cntk_model_dump)) {
        std::printf("[SIMD::CNN] Configuration file not found!\n");
        return -1;
    }
    Model m;
    if (!convert(d, m) || m.min_w <= 0 || m.min_h <= 0) {
        std::printf("[ccnn] %s: not a CNTK dump of a three-layer cascade network\n", cntk_model_dump);
        return -1;
    }
    return save(m, binary_file) ? 0 : -1;
}
