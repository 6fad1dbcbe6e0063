// pt_files.cpp -- the reference's on-disk formats: checkpoint / resume and
// the ASCII snapshots.
//
// print_restart (PS/main.f90:457-496) writes, for RaT != 0, comp = 1,
// tracer = 1 and iC_continuous = 1:
//   Trestart         one record: xg(-6:nx+6), zg(-6:nz+6), T(-3:nx+3,-3:nz+3),
//                    Cnew(1,-5:nx+5,-5:nz+5), SF(-3:nx+3,-3:nz+3)
//   restart_tracers  one record: rtr, str, Ctr, Ttr (ntr each, tracer ID order);
//                    with discrete types (iC_continuous = 0) rtr, str, Ttype
//                    (INTEGER*4), Ttr
// as Fortran unformatted sequential files: each record is framed by its
// length in bytes as a 4-byte little-endian integer before and after.  The
// restart=1 branch of build_grid (PS/grid.f90:55-103) reads them back.
// Records longer than 2^31-1 bytes (beyond 67 M tracers) would need the
// compilers' sub-record framing and are refused.
//
// snapshots (PS/main.f90:377-424) writes T1xxxx (per node, i outer, k inner:
// xg, zg, T, SF, u, w, 2*vis*strain, vis, Tvis as g14.5E3, nempty as i6,
// Cnew, Cvis as g14.5E3; a list-directed blank record after each i) and
// P1xxxx (per k: zg and the lateral means of T, strain and 2*vis*strain).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "fortran_format.hpp"
#include "pt_internal.hpp"

namespace pt {

namespace {

std::string join(const char* dir, const char* name) {
  std::string d = dir ? dir : ".";
  if (!d.empty() && d.back() != '/') d += '/';
  return d + name;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

// one record from the parts, framed by its length
// one record from byte spans (pointer, bytes)
int write_record(const std::string& path, const std::vector<std::pair<const void*, size_t>>& parts) {
  size_t bytes = 0;
  for (auto& p : parts) bytes += p.second;
  if (bytes > 0x7fffffffu) {
    set_error(path + ": record of " + std::to_string(bytes) + " bytes exceeds the 4-byte record length");
    return PT_ERR_UNSUPPORTED;
  }
  File fh;
  fh.f = std::fopen(path.c_str(), "wb");
  if (!fh.f) {
    set_error("cannot open " + path + " for writing");
    return PT_ERR_INVALID;
  }
  const uint32_t len = (uint32_t)bytes;
  bool ok = std::fwrite(&len, 4, 1, fh.f) == 1;
  for (auto& p : parts) ok = ok && std::fwrite(p.first, 1, p.second, fh.f) == p.second;
  ok = ok && std::fwrite(&len, 4, 1, fh.f) == 1;
  ok = (std::fclose(fh.f) == 0) && ok;
  fh.f = nullptr;
  if (!ok) {
    set_error("write to " + path + " failed");
    return PT_ERR_INVALID;
  }
  return PT_OK;
}

// the single record of `path`, which must hold exactly `bytes` bytes
int read_record_bytes(const std::string& path, size_t bytes, void* out) {
  File fh;
  fh.f = std::fopen(path.c_str(), "rb");
  if (!fh.f) {
    set_error("cannot open " + path);
    return PT_ERR_INVALID;
  }
  uint32_t head = 0, tail = 0;
  if (std::fread(&head, 4, 1, fh.f) != 1 || head != bytes) {
    set_error(path + ": record of " + std::to_string(head) + " bytes, this context expects " + std::to_string(bytes));
    return PT_ERR_INVALID;
  }
  if (std::fread(out, 1, bytes, fh.f) != bytes || std::fread(&tail, 4, 1, fh.f) != 1 || tail != head) {
    set_error(path + ": truncated or malformed record");
    return PT_ERR_INVALID;
  }
  return PT_OK;
}
// the single record of `path`, which must hold exactly `count` doubles
int read_record(const std::string& path, size_t count, std::vector<double>& out) {
  out.resize(count);
  return read_record_bytes(path, sizeof(double) * count, out.data());
}

// xg, zg of a uniform grid as build_grid forms them (PS/grid.f90:46-116):
// nodes r = dr*i, mirrored halos, exact end points
void grid_coords(const Tables& t, std::vector<double>& xg, std::vector<double>& zg) {
  const int h = 6;   // span + span_interp
  xg.assign(t.nx + 1 + 2 * h, 0.0);
  zg.assign(t.nz + 1 + 2 * h, 0.0);
  auto X = [&](int i) -> double& { return xg[i + h]; };
  auto Z = [&](int k) -> double& { return zg[k + h]; };
  for (int i = 0; i <= t.nx; i++) X(i) = t.dr * (double)i;
  for (int k = 0; k <= t.nz; k++) Z(k) = t.ds * (double)k;
  for (int i = 1; i <= h; i++) {
    X(-i) = -X(i);
    X(t.nx + i) = t.aspect - X(t.nx - i) + t.aspect;
  }
  for (int k = 1; k <= h; k++) {
    Z(-k) = -Z(k);
    Z(t.nz + k) = 1.0 - Z(t.nz - k) + 1.0;
  }
  X(0) = 0.0; Z(0) = 0.0;
  X(t.nx) = t.aspect; Z(t.nz) = 1.0;
}

// integrate1Dx (PS/basics.f90:494-515), xr = 1 on a uniform grid
double integrate1Dx(const double* f, int nx, double dr) {
  double fint = 0.0;
  for (int i = 0; i < nx; i++) fint = fint + 0.5 * (f[i] * 1.0 + f[i + 1] * 1.0) * dr;
  return fint;
}

}  // namespace

int write_snapshot(Context& c, const char* dir, int fcount) {
  const Tables& t = c.T;
  const int nx = t.nx, nz = t.nz;
  if (fcount < 0 || fcount > 89999) {
    set_error("snapshot number must be 0..89999 (file names T10000..T99999)");
    return PT_ERR_INVALID;
  }
  const size_t n3 = (size_t)(nx + 7) * (nz + 7), n5 = (size_t)(nx + 11) * (nz + 11), nn = (size_t)(nx + 1) * (nz + 1);
  std::vector<double> T(n3), SF(n3), u(n3), w(n3), vis(n3), Tvis(n3), C(n5), Cv(n5), st(nn);
  std::vector<int32_t> ne(nn);
  if (!c.strain_buf) PT_HIP(hipMalloc((void**)&c.strain_buf, sizeof(double) * nn));
  PT_TRY(strain_field(c, c.strain_buf));
  const double* Tdev = c.t_initial ? c.T0 : c.Tratio;   // the module T
  const std::pair<double*, const double*> f3[] = {{T.data(), Tdev}, {SF.data(), c.SF}, {u.data(), c.u},
                                                  {w.data(), c.w}, {vis.data(), c.L[1].v.vis}, {Tvis.data(), c.Tvis}};
  for (auto& q : f3) PT_HIP(hipMemcpyAsync(q.first, q.second, sizeof(double) * n3, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(C.data(), c.Cnew, sizeof(double) * n5, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(Cv.data(), c.Cvis, sizeof(double) * n5, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(st.data(), c.strain_buf, sizeof(double) * nn, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(ne.data(), c.nempty, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> xg, zg;
  grid_coords(t, xg, zg);
  auto I3 = [&](int i, int k) { return (size_t)(k + 3) * (nx + 7) + (i + 3); };
  auto I5 = [&](int i, int k) { return (size_t)(k + 5) * (nx + 11) + (i + 5); };
  auto IN = [&](int i, int k) { return (size_t)k * (nx + 1) + i; };
  char name[8];
  std::snprintf(name, sizeof name, "T%5d", 10000 + fcount);
  std::string body;
  body.reserve((size_t)(nx + 1) * ((nz + 1) * 161 + 3));
  for (int i = 0; i <= nx; i++) {
    for (int k = 0; k <= nz; k++) {
      const double stress = 2.0 * vis[I3(i, k)] * st[IN(i, k)];
      const double v[9] = {xg[i + 6], zg[k + 6], T[I3(i, k)], SF[I3(i, k)], u[I3(i, k)], w[I3(i, k)],
                           stress, vis[I3(i, k)], Tvis[I3(i, k)]};
      for (double x : v) g14_5e3(x, body);
      char nb[16];
      std::snprintf(nb, sizeof nb, "%6d", ne[IN(i, k)]);
      body += nb;
      g14_5e3(C[I5(i, k)], body);
      g14_5e3(Cv[I5(i, k)], body);
      body += '\n';
    }
    body += "  \n";   // write(101,*) " "
  }
  std::string prof;
  std::vector<double> rt(nx + 1), rs(nx + 1), rq(nx + 1);
  for (int k = 0; k <= nz; k++) {
    for (int i = 0; i <= nx; i++) {
      rt[i] = T[I3(i, k)];
      rs[i] = st[IN(i, k)];
      rq[i] = 2.0 * vis[I3(i, k)] * st[IN(i, k)];
    }
    g14_5e3(zg[k + 6], prof);
    g14_5e3(integrate1Dx(rt.data(), nx, t.dr) / t.aspect, prof);
    g14_5e3(integrate1Dx(rs.data(), nx, t.dr) / t.aspect, prof);
    g14_5e3(integrate1Dx(rq.data(), nx, t.dr) / t.aspect, prof);
    prof += '\n';
  }
  for (int pass = 0; pass < 2; pass++) {
    std::string path = join(dir, name);
    if (pass == 1) path[path.size() - 6] = 'P';
    const std::string& text = pass == 0 ? body : prof;
    File fh;
    fh.f = std::fopen(path.c_str(), "w");
    if (!fh.f) {
      set_error("cannot open " + path + " for writing");
      return PT_ERR_INVALID;
    }
    bool ok = std::fwrite(text.data(), 1, text.size(), fh.f) == text.size();
    ok = (std::fclose(fh.f) == 0) && ok;
    fh.f = nullptr;
    if (!ok) {
      set_error("write to " + path + " failed");
      return PT_ERR_INVALID;
    }
  }
  return PT_OK;
}

int write_restart(Context& c, const char* dir) {
  const Tables& t = c.T;
  const size_t n3 = (size_t)(t.nx + 7) * (t.nz + 7), n5 = (size_t)(t.nx + 11) * (t.nz + 11);
  const size_t ntr = (size_t)t.ntr;
  std::vector<double> xg, zg, T(n3), C(n5), SF(n3), tr(4 * ntr);
  grid_coords(t, xg, zg);
  // the module T: the initial field until the first energy_time, then Tratio (PS/energy.f90:552)
  const double* Tdev = c.t_initial ? c.T0 : c.Tratio;
  PT_HIP(hipMemcpyAsync(T.data(), Tdev, sizeof(double) * n3, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(C.data(), c.Cnew, sizeof(double) * n5, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(SF.data(), c.SF, sizeof(double) * n3, hipMemcpyDeviceToHost, c.stream));
  double *dr = nullptr, *ds = nullptr, *dT = nullptr, *dC = nullptr;
  PT_TRY(tracers_in_id_order(c, &dr, &ds, &dT, &dC));
  const size_t tb = sizeof(double) * ntr;
  PT_HIP(hipMemcpyAsync(tr.data(), dr, tb, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(tr.data() + ntr, ds, tb, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(tr.data() + 2 * ntr, dC, tb, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipMemcpyAsync(tr.data() + 3 * ntr, dT, tb, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  const size_t d = sizeof(double);
  PT_TRY(write_record(join(dir, "Trestart"), {{xg.data(), d * xg.size()},
                                              {zg.data(), d * zg.size()},
                                              {T.data(), d * n3},
                                              {C.data(), d * n5},
                                              {SF.data(), d * n3}}));
  if (c.P.iC_continuous == 1)
    return write_record(join(dir, "restart_tracers"), {{tr.data(), d * tr.size()}});
  // discrete types: rtr, str, Ttype (INTEGER*4), Ttr (PS/main.f90:488-489)
  std::vector<int32_t> type(ntr);
  for (size_t j = 0; j < ntr; j++) type[j] = (int32_t)tr[2 * ntr + j];
  return write_record(join(dir, "restart_tracers"),
                      {{tr.data(), 2 * tb}, {type.data(), sizeof(int32_t) * ntr}, {tr.data() + 3 * ntr, tb}});
}

int read_restart(Context& c, const char* dir, double Tbot, double Ttop) {
  const Tables& t = c.T;
  const size_t n3 = (size_t)(t.nx + 7) * (t.nz + 7), n5 = (size_t)(t.nx + 11) * (t.nz + 11);
  const size_t ntr = (size_t)t.ntr;
  std::vector<double> xg, zg, rec, tr;
  grid_coords(t, xg, zg);
  PT_TRY(read_record(join(dir, "Trestart"), xg.size() + zg.size() + 2 * n3 + n5, rec));
  // the file's grid must be this context's uniform grid
  double dev = 0.0;
  for (size_t i = 0; i < xg.size(); i++) dev = std::max(dev, std::fabs(rec[i] - xg[i]));
  for (size_t k = 0; k < zg.size(); k++) dev = std::max(dev, std::fabs(rec[xg.size() + k] - zg[k]));
  if (!(dev <= 1e-12 * std::max(1.0, t.aspect))) {
    set_error("Trestart grid coordinates differ from this context's grid by " + std::to_string(dev));
    return PT_ERR_INVALID;
  }
  if (c.P.iC_continuous == 1) {
    PT_TRY(read_record(join(dir, "restart_tracers"), 4 * ntr, tr));
  } else {
    // rtr, str, Ttype (INTEGER*4), Ttr (PS/grid.f90:86-87)
    std::vector<char> raw(28 * ntr);
    PT_TRY(read_record_bytes(join(dir, "restart_tracers"), raw.size(), raw.data()));
    tr.resize(4 * ntr);
    std::memcpy(tr.data(), raw.data(), 16 * ntr);
    for (size_t j = 0; j < ntr; j++) {
      int32_t ty;
      std::memcpy(&ty, raw.data() + 16 * ntr + 4 * j, 4);
      if (ty != 0 && ty != 1) {
        set_error("restart_tracers: tracer type " + std::to_string(ty) + " outside {0,1}");
        return PT_ERR_INVALID;
      }
      tr[2 * ntr + j] = (double)ty;
    }
    std::memcpy(tr.data() + 3 * ntr, raw.data() + 20 * ntr, 8 * ntr);
  }
  const double* T3 = rec.data() + xg.size() + zg.size();
  const double* C5 = T3 + n3;
  const double* SF3 = C5 + n5;
  const double *r = tr.data(), *s = r + ntr, *Cq = s + ntr, *Tq = Cq + ntr;
  return restart_state(c, T3, C5, SF3, Tbot, Ttop, r, s, Tq, Cq);
}

}  // namespace pt
