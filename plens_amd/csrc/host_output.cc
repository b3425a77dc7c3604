// Data output files of PLENS::write() (dgsem/plens.cc:4170-4322) for one process: the VTU piece
// with one VTK Lagrange hexahedron per cell, its PVTU record, the PVD collection and the .times
// log. The fields come from plens_gpu_output_fields; coordinates from plens_mesh_output_points.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "plens_host.h"

namespace {

const char* const kNodalNames[15] = {"txx", "txy", "txz", "tyy", "tyz", "tzz", "qx", "qy", "qz",
                                     "rho", "p", "T", "u", "v", "w"};

std::string b64(const void* data, size_t n) {
    static const char* tbl = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    const unsigned char* b = static_cast<const unsigned char*>(data);
    std::string out;
    out.reserve(4 * ((n + 2) / 3));
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const uint32_t v = (uint32_t)b[i] << 16 | (uint32_t)b[i + 1] << 8 | b[i + 2];
        out += tbl[v >> 18]; out += tbl[(v >> 12) & 63]; out += tbl[(v >> 6) & 63]; out += tbl[v & 63];
    }
    if (i < n) {
        uint32_t v = (uint32_t)b[i] << 16;
        if (i + 1 < n) v |= (uint32_t)b[i + 1] << 8;
        out += tbl[v >> 18];
        out += tbl[(v >> 12) & 63];
        out += (i + 1 < n) ? tbl[(v >> 6) & 63] : '=';
        out += '=';
    }
    return out;
}

// inline binary DataArray body: UInt64 byte count and the bytes, each base64-encoded on its own
std::string vtk_binary(const void* data, size_t bytes) {
    const uint64_t hdr = bytes;
    return b64(&hdr, sizeof(hdr)) + b64(data, bytes);
}

// position of lattice point (i, j, k) of an order-n Lagrange hexahedron in VTK's point order
// (vtkHigherOrderHexahedron::PointIndexFromIJK, file version 2.2: corners, edges, faces, body)
int vtk_hex_index(int i, int j, int k, int n) {
    const bool ib = i == 0 || i == n, jb = j == 0 || j == n, kb = k == 0 || k == n;
    const int nb = ib + jb + kb;
    if (nb == 3) return (i ? (j ? 2 : 1) : (j ? 3 : 0)) + (k ? 4 : 0);
    int off = 8;
    const int m = n - 1;
    if (nb == 2) {
        if (!ib) return (i - 1) + (j ? 2 * m : 0) + (k ? 4 * m : 0) + off;
        if (!jb) return (j - 1) + (i ? m : 3 * m) + (k ? 4 * m : 0) + off;
        off += 8 * m;
        return (k - 1) + m * (i ? (j ? 2 : 1) : (j ? 3 : 0)) + off;
    }
    off += 12 * m;
    if (nb == 1) {
        if (ib) return (j - 1) + m * (k - 1) + (i ? m * m : 0) + off;
        off += 2 * m * m;
        if (jb) return (i - 1) + m * (k - 1) + (j ? m * m : 0) + off;
        off += 2 * m * m;
        return (i - 1) + m * (j - 1) + (k ? m * m : 0) + off;
    }
    off += 6 * m * m;
    return off + (i - 1) + m * ((j - 1) + m * (k - 1));
}

std::string counter_str(int32_t c) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%06d", (int)c);
    return buf;
}

bool write_file(const std::string& path, const std::string& text, const char* mode = "wb") {
    FILE* f = std::fopen(path.c_str(), mode);
    if (!f) return false;
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace

extern "C" int plens_output_write(const plens_mesh* m, const char* directory, const char* base, int32_t counter,
                                  double time, double wall_time, const double* nodal, const double* cell,
                                  const double* record_times, int32_t n_records) {
    if (!m || !directory || !base || !nodal || !cell || counter < 0 || n_records < 0 ||
        (n_records > 0 && !record_times))
        return PLENS_E_INVALID;
    const int N = plens_mesh_degree(m), nc = plens_mesh_n_cells(m);
    if (N < 1 || nc < 1) return PLENS_E_INVALID;
    const int np = N + 1, n = np * np * np;
    const size_t npts = (size_t)nc * n;
    std::vector<double> X(3 * npts);
    if (plens_mesh_output_points(m, X.data())) return PLENS_E_INVALID;
    // connectivity: VTK order -> lexicographic lattice point of the cell
    std::vector<int> perm(n);
    for (int k = 0; k < np; ++k)
        for (int j = 0; j < np; ++j)
            for (int i = 0; i < np; ++i) perm[vtk_hex_index(i, j, k, N)] = i + np * (j + np * k);
    std::vector<int64_t> conn(npts), offs(nc);
    std::vector<uint8_t> types(nc, 72);
    for (int c = 0; c < nc; ++c) {
        for (int v = 0; v < n; ++v) conn[(size_t)c * n + v] = (int64_t)c * n + perm[v];
        offs[c] = (int64_t)(c + 1) * n;
    }
    const std::string stem = std::string(base) + "_" + counter_str(counter);
    const std::string dir = std::string(directory) + "/";
    std::string x;
    x += "<?xml version=\"1.0\"?>\n";
    x += "<VTKFile type=\"UnstructuredGrid\" version=\"2.2\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    x += "<UnstructuredGrid>\n";
    x += "<Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" + std::to_string(nc) + "\">\n";
    x += "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"binary\">\n";
    x += vtk_binary(X.data(), sizeof(double) * X.size()) + "\n</DataArray>\n</Points>\n";
    x += "<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" format=\"binary\">\n";
    x += vtk_binary(conn.data(), sizeof(int64_t) * conn.size()) + "\n</DataArray>\n";
    x += "<DataArray type=\"Int64\" Name=\"offsets\" format=\"binary\">\n";
    x += vtk_binary(offs.data(), sizeof(int64_t) * offs.size()) + "\n</DataArray>\n";
    x += "<DataArray type=\"UInt8\" Name=\"types\" format=\"binary\">\n";
    x += vtk_binary(types.data(), types.size()) + "\n</DataArray>\n</Cells>\n";
    x += "<PointData Scalars=\"rho\">\n";
    std::vector<std::string> names(kNodalNames, kNodalNames + 15);
    for (int f = 0; f < 15; ++f) {
        x += std::string("<DataArray type=\"Float64\" Name=\"") + kNodalNames[f] + "\" format=\"binary\">\n";
        x += vtk_binary(nodal + (size_t)f * npts, sizeof(double) * npts) + "\n</DataArray>\n";
    }
    // cell quantities, constant over the cell's points: subdomain, alpha, alpha_d, loc_dt, index
    const char* cnames[5] = {"subdomain", "alpha", "alpha_d", "loc_dt", "global_active_cell_index"};
    std::vector<double> tmp(npts);
    for (int f = 0; f < 5; ++f) {
        for (int c = 0; c < nc; ++c) {
            double v = 0;
            if (f == 0) v = 0;
            else if (f == 4) v = c;
            else v = cell[(size_t)(f - 1) * nc + c];
            for (int i = 0; i < n; ++i) tmp[(size_t)c * n + i] = v;
        }
        x += std::string("<DataArray type=\"Float64\" Name=\"") + cnames[f] + "\" format=\"binary\">\n";
        x += vtk_binary(tmp.data(), sizeof(double) * npts) + "\n</DataArray>\n";
        names.push_back(cnames[f]);
    }
    x += "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    if (!write_file(dir + stem + ".0.vtu", x)) return PLENS_E_INVALID;

    std::string pv;
    pv += "<?xml version=\"1.0\"?>\n";
    pv += "<VTKFile type=\"PUnstructuredGrid\" version=\"2.2\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    pv += "<PUnstructuredGrid GhostLevel=\"0\">\n<PPointData Scalars=\"rho\">\n";
    for (const auto& nm : names) pv += "<PDataArray type=\"Float64\" Name=\"" + nm + "\"/>\n";
    pv += "</PPointData>\n<PPoints>\n<PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n</PPoints>\n";
    pv += "<Piece Source=\"" + stem + ".0.vtu\"/>\n</PUnstructuredGrid>\n</VTKFile>\n";
    if (!write_file(dir + stem + ".pvtu", pv)) return PLENS_E_INVALID;

    std::string pd;
    pd += "<?xml version=\"1.0\"?>\n";
    pd += "<VTKFile type=\"Collection\" version=\"0.1\" ByteOrder=\"LittleEndian\">\n<Collection>\n";
    char buf[64];
    for (int r = 0; r < n_records; ++r) {
        std::snprintf(buf, sizeof(buf), "%.17g", record_times[r]);
        pd += std::string("<DataSet timestep=\"") + buf + "\" group=\"\" part=\"0\" file=\"" + base + "_" +
              counter_str(r) + ".pvtu\"/>\n";
    }
    pd += "</Collection>\n</VTKFile>\n";
    if (!write_file(dir + base + ".pvd", pd)) return PLENS_E_INVALID;

    std::snprintf(buf, sizeof(buf), "%.17g", time);
    std::string line = std::to_string(counter) + " " + buf;
    std::snprintf(buf, sizeof(buf), "%.17g", wall_time);
    line += std::string(" ") + buf + "\n";
    if (!write_file(dir + base + ".times", line, "ab")) return PLENS_E_INVALID;
    return PLENS_OK;
}
