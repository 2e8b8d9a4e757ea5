// On-disk formats written from device-resident fields.
//
// .vti: VTK XML image data with raw appended cell arrays, the layout of
//   R/src/IO/ExportVTK.jl:16-111 (density, L1_Stress, VonMises, optionally
//   PrincipalValues and the max/min principal directions; each array is a
//   UInt32 byte count followed by Float32 data).
// .bin: the checkpoint/web format of R/src/IO/Postprocessing.jl:370-471
//   ("HEXA" magic, version 2, iteration, radius, threshold, count, dx dy dz,
//   centroids, densities, signed l1, config JSON), read back by
//   R/src/IO/Configuration.jl:32-78.  The reference writes only elements with
//   rho >= threshold (optionally strided) while its reader takes count as the
//   full element count, so a restart fails unless every element survives;
//   full_field = 1 writes every element in order, which that reader restarts
//   from.
//
// The stress fields come from the stress kernel applied to the handle's
// device solution and density; only the finished arrays cross PCIe, once.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "handle.h"
#include "hexa_internal.h"

using namespace hexa;

namespace {

// Julia's print(::Float32): shortest round-trip digits, plain decimal for
// 1e-4 <= |x| < 1e6 (always with a fractional part), otherwise d.ddde[-]n.
std::string julia_f32(float x) {
    if (x == 0.0f) return std::signbit(x) ? "-0.0" : "0.0";
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "Inf" : "-Inf";
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::scientific);
    std::string s(buf, res.ptr);
    std::string sign;
    if (s[0] == '-') { sign = "-"; s = s.substr(1); }
    const size_t epos = s.find('e');
    std::string mant = s.substr(0, epos);
    const int e10 = std::stoi(s.substr(epos + 1));
    std::string digits;
    for (char c : mant)
        if (c != '.') digits += c;
    if (e10 >= -5 + 1 && e10 < 6) {
        std::string out;
        if (e10 >= 0) {
            if ((int)digits.size() <= e10 + 1) {
                out = digits + std::string(e10 + 1 - digits.size(), '0') + ".0";
            } else {
                out = digits.substr(0, e10 + 1) + "." + digits.substr(e10 + 1);
            }
        } else {
            out = "0." + std::string(-e10 - 1, '0') + digits;
        }
        return sign + out;
    }
    std::string m = digits.substr(0, 1) + "." + (digits.size() > 1 ? digits.substr(1) : std::string("0"));
    return sign + m + "e" + std::to_string(e10);
}

struct Fields {
    std::vector<float> density, l1, vm, principal, dmax, dmin;
};

// Density (given, or the handle's as Float32) and stress fields on the host.
int32_t gather_fields(hexa_handle* h, const float* density, double E, bool vectors, Fields& F) {
    if (!h->have_solution) { set_error("no solution on the device: solve first"); return HEXA_E_INVALID; }
    if (!density && !h->have_density) { set_error("no density"); return HEXA_E_INVALID; }
    const int64_t ne = h->g.ne;
    hipStream_t st = h->stream;
    HX_TRY(dalloc(h, &h->d_stress, (size_t)(17 * ne)));
    float* base = h->d_stress;
    float* dp = base;
    float* dv = base + 3 * ne;
    float* dl = base + 4 * ne;
    float* dmx = base + 5 * ne;
    float* dmn = base + 8 * ne;
    launch_stress(st, h->g, h->grid.dx, h->grid.dy, h->grid.dz, E, h->opt.nu, h->d_x, h->d_rho, dp, dv, dl,
                  vectors ? dmx : nullptr, vectors ? dmn : nullptr, nullptr);
    HX_TRY(check_launch());
    F.l1.resize(ne);
    F.vm.resize(ne);
    F.principal.resize(3 * ne);
    HX_CHECK(hipMemcpyAsync(F.l1.data(), dl, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipMemcpyAsync(F.vm.data(), dv, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
    HX_CHECK(hipMemcpyAsync(F.principal.data(), dp, sizeof(float) * 3 * ne, hipMemcpyDeviceToHost, st));
    if (vectors) {
        F.dmax.resize(3 * ne);
        F.dmin.resize(3 * ne);
        HX_CHECK(hipMemcpyAsync(F.dmax.data(), dmx, sizeof(float) * 3 * ne, hipMemcpyDeviceToHost, st));
        HX_CHECK(hipMemcpyAsync(F.dmin.data(), dmn, sizeof(float) * 3 * ne, hipMemcpyDeviceToHost, st));
    }
    F.density.resize(ne);
    if (density) {
        std::memcpy(F.density.data(), density, sizeof(float) * ne);
    } else {
        HX_TRY(dalloc(h, &h->d_f0, (size_t)ne));
        launch_convert<double, float>(st, ne, h->d_rho, h->d_f0);
        HX_TRY(check_launch());
        HX_CHECK(hipMemcpyAsync(F.density.data(), h->d_f0, sizeof(float) * ne, hipMemcpyDeviceToHost, st));
    }
    HX_CHECK(hipStreamSynchronize(st));
    return HEXA_OK;
}

bool put(FILE* f, const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }

}  // namespace

extern "C" {

int32_t hexa_write_vti(hexa_handle* h, const char* path, const float* density, double E, double ox, double oy,
                       double oz, int32_t write_vectors) {
    if (!h || !path) { set_error("null argument"); return HEXA_E_INVALID; }
    Fields F;
    HX_TRY(gather_fields(h, density, E, write_vectors != 0, F));
    std::string fn(path);
    if (fn.size() < 4 || fn.compare(fn.size() - 4, 4, ".vti") != 0) fn += ".vti";
    const int64_t ne = h->g.ne;
    const std::string ext = "0 " + std::to_string(h->g.nx) + " 0 " + std::to_string(h->g.ny) + " 0 " +
                            std::to_string(h->g.nz);
    std::string x = "<?xml version=\"1.0\"?>\n"
                    "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    x += "  <ImageData WholeExtent=\"" + ext + "\" Origin=\"" + julia_f32((float)ox) + " " + julia_f32((float)oy) +
         " " + julia_f32((float)oz) + "\" Spacing=\"" + julia_f32((float)h->grid.dx) + " " +
         julia_f32((float)h->grid.dy) + " " + julia_f32((float)h->grid.dz) + "\">\n";
    x += "    <Piece Extent=\"" + ext + "\">\n";
    x += "      <CellData Scalars=\"Density\" Vectors=\"MaxPrincipalDirection\">\n";
    int64_t off = 0;
    auto header = [&](const char* name, int comps, int64_t n) {
        x += std::string("        <DataArray type=\"Float32\" Name=\"") + name + "\"" +
             (comps > 1 ? " NumberOfComponents=\"" + std::to_string(comps) + "\"" : std::string()) +
             " format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
        off += 4 + n * 4;
    };
    header("Density", 1, ne);
    header("L1_Stress", 1, ne);
    header("VonMises", 1, ne);
    if (write_vectors) {
        header("PrincipalValues", 3, 3 * ne);
        header("MaxPrincipalDirection", 3, 3 * ne);
        header("MinPrincipalDirection", 3, 3 * ne);
    }
    x += "      </CellData>\n    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n_";
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f) { set_error("cannot open %s for writing", fn.c_str()); return HEXA_E_INVALID; }
    bool ok = put(f, x.data(), x.size());
    auto arr = [&](const std::vector<float>& a) {
        const uint32_t nb = (uint32_t)(a.size() * sizeof(float));
        ok = ok && put(f, &nb, 4) && put(f, a.data(), nb);
    };
    arr(F.density);
    arr(F.l1);
    arr(F.vm);
    if (write_vectors) {
        arr(F.principal);
        arr(F.dmax);
        arr(F.dmin);
    }
    const char* tail = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && put(f, tail, std::strlen(tail));
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { set_error("write to %s failed", fn.c_str()); return HEXA_E_INVALID; }
    return HEXA_OK;
}

int32_t hexa_write_bin(hexa_handle* h, const char* path, const float* density, double E, int32_t iter,
                       float radius, float threshold, int64_t max_export_cells, int32_t full_field,
                       const char* json) {
    if (!h || !path) { set_error("null argument"); return HEXA_E_INVALID; }
    Fields F;
    HX_TRY(gather_fields(h, density, E, false, F));
    const int64_t ne = h->g.ne;
    // element selection (Postprocessing.jl:383-402)
    std::vector<int64_t> idx;
    if (full_field) {
        idx.resize(ne);
        for (int64_t e = 0; e < ne; ++e) idx[e] = e;
    } else {
        std::vector<int64_t> act;
        for (int64_t e = 0; e < ne; ++e)
            if (F.density[e] >= threshold) act.push_back(e);
        const int64_t na = (int64_t)act.size();
        if (na == 0) return HEXA_OK;   // the reference writes nothing (:384)
        if (max_export_cells > 0 && na > max_export_cells) {
            const double step = (double)na / (double)max_export_cells;
            for (double cur = 1.0; cur <= (double)na; cur += step) {
                const int64_t k = (int64_t)std::floor(cur);
                if (k <= na) idx.push_back(act[k - 1]);
            }
        } else {
            idx = act;
        }
    }
    const int64_t cnt = (int64_t)idx.size();
    std::vector<float> cen(3 * cnt), dens(cnt), sl1(cnt);
    const int nx = h->g.nx, ny = h->g.ny;
    const float hx = (float)h->grid.dx, hy = (float)h->grid.dy, hz = (float)h->grid.dz;
    static const int CX[8] = {0, 1, 1, 0, 0, 1, 1, 0}, CY[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                     CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int64_t q = 0; q < cnt; ++q) {
        const int64_t e = idx[q];
        const int i = (int)(e % nx), j = (int)((e / nx) % ny), k = (int)(e / ((int64_t)nx * ny));
        // centroid = mean of the 8 Float32 node coordinates, in node order
        float s[3] = {0.f, 0.f, 0.f};
        for (int a = 0; a < 8; ++a) {
            s[0] += (float)(i + CX[a]) * hx;
            s[1] += (float)(j + CY[a]) * hy;
            s[2] += (float)(k + CZ[a]) * hz;
        }
        for (int c = 0; c < 3; ++c) cen[3 * q + c] = s[c] / 8.0f;
        dens[q] = F.density[e];
        // l1 signed by the principal value of largest magnitude (:446-452)
        const float p1 = F.principal[3 * e], p2 = F.principal[3 * e + 1], p3 = F.principal[3 * e + 2];
        float am = std::fabs(p1), sv = p1 > 0 ? 1.f : (p1 < 0 ? -1.f : 0.f);
        if (std::fabs(p2) > am) { am = std::fabs(p2); sv = p2 > 0 ? 1.f : (p2 < 0 ? -1.f : 0.f); }
        if (std::fabs(p3) > am) { am = std::fabs(p3); sv = p3 > 0 ? 1.f : (p3 < 0 ? -1.f : 0.f); }
        sl1[q] = F.l1[e] * sv;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) { set_error("cannot open %s for writing", path); return HEXA_E_INVALID; }
    const uint32_t magic = 0x48455841u, version = 2u, ucount = (uint32_t)cnt;
    const int32_t it = iter;
    bool ok = put(f, &magic, 4) && put(f, &version, 4) && put(f, &it, 4) && put(f, &radius, 4) &&
              put(f, &threshold, 4) && put(f, &ucount, 4) && put(f, &hx, 4) && put(f, &hy, 4) && put(f, &hz, 4);
    ok = ok && put(f, cen.data(), sizeof(float) * 3 * cnt) && put(f, dens.data(), sizeof(float) * cnt) &&
         put(f, sl1.data(), sizeof(float) * cnt);
    const std::string js = json ? json : "{}";
    const uint32_t jl = (uint32_t)js.size();
    ok = ok && put(f, &jl, 4) && put(f, js.data(), js.size());
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { set_error("write to %s failed", path); return HEXA_E_INVALID; }
    return HEXA_OK;
}

int32_t hexa_regrid_density(const hexa_grid* old_grid, const float* d_old_density, const float* d_old_alpha,
                            const hexa_grid* new_grid, double len_x, double len_y, double len_z, int32_t hard,
                            float* d_new_density, float* d_new_alpha, void* stream) {
    if (!old_grid || !new_grid || !d_old_density || !d_new_density) { set_error("null argument"); return HEXA_E_INVALID; }
    if (old_grid->nx < 1 || old_grid->ny < 1 || old_grid->nz < 1 || new_grid->nx < 1 || new_grid->ny < 1 ||
        new_grid->nz < 1 || !(len_x > 0) || !(len_y > 0) || !(len_z > 0)) {
        set_error("invalid grids or domain lengths");
        return HEXA_E_INVALID;
    }
    hipStream_t s = (hipStream_t)stream;
    launch_regrid(s, old_grid->nx, old_grid->ny, old_grid->nz, new_grid->nx, new_grid->ny, new_grid->nz,
                  (float)new_grid->dx, (float)new_grid->dy, (float)new_grid->dz, (float)len_x, (float)len_y,
                  (float)len_z, hard, d_old_density, d_old_alpha, d_new_density, d_new_alpha);
    HX_TRY(check_launch());
    HX_CHECK(hipStreamSynchronize(s));
    return HEXA_OK;
}

int32_t hexa_regrid_density_host(const hexa_grid* old_grid, const float* old_density, const float* old_alpha,
                                 const hexa_grid* new_grid, double len_x, double len_y, double len_z, int32_t hard,
                                 float* new_density, float* new_alpha) {
    if (!old_grid || !new_grid || !old_density || !new_density) { set_error("null argument"); return HEXA_E_INVALID; }
    const size_t no = (size_t)old_grid->nx * old_grid->ny * old_grid->nz;
    const size_t nn = (size_t)new_grid->nx * new_grid->ny * new_grid->nz;
    float *d_o = nullptr, *d_oa = nullptr, *d_n = nullptr, *d_na = nullptr;
    hipStream_t s = nullptr;   // a stream of its own: nothing here touches the legacy stream
    int32_t st = HEXA_OK;
    {
        DeviceOpGuard lock;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            set_error("stream creation failed");
            st = HEXA_E_DEVICE;
        }
        auto alloc = [&](float** p, size_t n) {
            if (st == HEXA_OK && hipMalloc((void**)p, sizeof(float) * (n ? n : 1)) != hipSuccess) {
                set_error("device allocation failed");
                st = HEXA_E_OOM;
            }
        };
        alloc(&d_o, no);
        alloc(&d_n, nn);
        if (old_alpha && new_alpha) { alloc(&d_oa, no); alloc(&d_na, nn); }
    }
    auto copy = [&](void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
        if (st == HEXA_OK && hipMemcpyAsync(dst, src, bytes, kind, s) != hipSuccess) st = HEXA_E_DEVICE;
    };
    copy(d_o, old_density, sizeof(float) * no, hipMemcpyHostToDevice);
    if (d_oa) copy(d_oa, old_alpha, sizeof(float) * no, hipMemcpyHostToDevice);
    if (st == HEXA_OK)
        st = hexa_regrid_density(old_grid, d_o, d_oa, new_grid, len_x, len_y, len_z, hard, d_n, d_na, s);
    copy(new_density, d_n, sizeof(float) * nn, hipMemcpyDeviceToHost);
    if (d_na) copy(new_alpha, d_na, sizeof(float) * nn, hipMemcpyDeviceToHost);
    if (st == HEXA_OK && hipStreamSynchronize(s) != hipSuccess) st = HEXA_E_DEVICE;
    if (st == HEXA_E_DEVICE) set_error("device copy failed");
    if (s) hipStreamSynchronize(s);
    DeviceOpGuard lock;
    hipFree(d_o); hipFree(d_oa); hipFree(d_n); hipFree(d_na);
    if (s) hipStreamDestroy(s);
    return st;
}

}  // extern "C"
