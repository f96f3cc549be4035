// state_io.hip -- the restart file (save_state / load_state, reference
// src/solver.f90:160-329): the interiors of p, u, v, dv_o%x, dv_o%y, then in
// 3-D w, dv_o%z, then in the two-phase mode vof, rho, mu, p_o, each as raw
// float64 in Fortran order, one after the other.  Host code over the
// library's own copies; the fields stay resident.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fen_internal.hpp"

namespace {

struct Slot {
    int field;      // FEN_* id, or -1 for p_o (two-phase mode, device only)
    bool ghosted;   // stored with ghosts on the device
};

std::vector<Slot> state_slots(const fen_ctx *c)
{
    std::vector<Slot> s = {{FEN_P, true}, {FEN_U, true}, {FEN_V, true}, {FEN_DVO_X, false}, {FEN_DVO_Y, false}};
    if (c->dim == 3) {
        s.push_back({FEN_W, true});
        s.push_back({FEN_DVO_Z, false});
    }
    if (c->mf) {
        s.push_back({FEN_VOF, true});
        s.push_back({FEN_RHO, true});
        s.push_back({FEN_MU, true});
        s.push_back({-1, true});
    }
    return s;
}

// interior (nx ny nz, i fastest) <-> ghosted (nx+2)(ny+2)(nz+2) host arrays;
// in 2-D the live plane is k = 1
void gather(const fen_ctx *c, const double *g, double *in)
{
    const long gx = c->nx + 2, gy = c->ny + 2;
    size_t q = 0;
    for (int k = 1; k <= c->nz; k++)
        for (int j = 1; j <= c->ny; j++) {
            const double *row = g + 1 + gx * (j + gy * k);
            std::memcpy(in + q, row, c->nx * sizeof(double));
            q += c->nx;
        }
}

void scatter(const fen_ctx *c, const double *in, double *g)
{
    const long gx = c->nx + 2, gy = c->ny + 2;
    size_t q = 0;
    for (int k = 1; k <= c->nz; k++)
        for (int j = 1; j <= c->ny; j++) {
            std::memcpy(g + 1 + gx * (j + gy * k), in + q, c->nx * sizeof(double));
            q += c->nx;
        }
}

int check(fen_ctx *c, int rc, const char *what)
{
    if (rc != FEN_OK && c->err.empty()) c->err = what;
    return rc;
}

}  // namespace

int fen_save_state(fen_ctx *c, const char *path)
{
    if (!c || !path) return FEN_ERR_ARG;
    const size_t ni = c->n_int;
    std::vector<double> g(c->n_ghost), in(ni);
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        c->err = std::string("cannot write ") + path;
        return FEN_ERR_ARG;
    }
    int rc = FEN_OK;
    for (const Slot &s : state_slots(c)) {
        if (s.field < 0) {   // p_o: FP64, device only
            if (hipMemcpy(g.data(), c->mf->p_o, c->n_ghost * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                c->err = "copy of p_o failed";
                rc = FEN_ERR_HIP;
                break;
            }
        } else if (s.ghosted) {
            rc = fen_download(c, s.field, g.data(), c->n_ghost);
        } else {
            rc = fen_download(c, s.field, in.data(), ni);
        }
        if (rc != FEN_OK) break;
        if (s.ghosted) gather(c, g.data(), in.data());
        if (std::fwrite(in.data(), sizeof(double), ni, fp) != ni) {
            c->err = std::string("short write to ") + path;
            rc = FEN_ERR_ARG;
            break;
        }
    }
    if (std::fclose(fp) != 0 && rc == FEN_OK) {
        c->err = std::string("cannot close ") + path;
        rc = FEN_ERR_ARG;
    }
    return rc;
}

int fen_load_state(fen_ctx *c, const char *path)
{
    if (!c || !path) return FEN_ERR_ARG;
    const std::vector<Slot> slots = state_slots(c);
    const size_t ni = c->n_int;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        c->err = std::string("cannot open ") + path;
        return FEN_ERR_ARG;
    }
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    const long want = (long)(slots.size() * ni * sizeof(double));
    if (size != want) {
        std::fclose(fp);
        c->err = std::string(path) + " holds " + std::to_string(size) + " bytes; this grid's state is " +
                 std::to_string(want);
        return FEN_ERR_ARG;
    }
    std::vector<double> g(c->n_ghost, 0.0), in(ni);
    int rc = FEN_OK;
    for (const Slot &s : slots) {
        if (std::fread(in.data(), sizeof(double), ni, fp) != ni) {
            c->err = std::string("short read from ") + path;
            rc = FEN_ERR_ARG;
            break;
        }
        if (!s.ghosted) {
            rc = check(c, fen_upload(c, s.field, in.data(), ni), "upload failed");
        } else {
            scatter(c, in.data(), g.data());   // ghosts follow from the fill below
            if (s.field == FEN_RHO || s.field == FEN_MU || s.field < 0) {
                double *dst = s.field == FEN_RHO ? c->mf->rho : (s.field == FEN_MU ? c->mf->mu : c->mf->p_o);
                if (hipMemcpy(dst, g.data(), c->n_ghost * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
                    c->err = "copy to the two-phase fields failed";
                    rc = FEN_ERR_HIP;
                }
            } else {
                rc = check(c, fen_upload(c, s.field, g.data(), c->n_ghost), "upload failed");
            }
        }
        if (rc != FEN_OK) break;
    }
    std::fclose(fp);
    if (rc != FEN_OK) return rc;
    // p%update_ghost_nodes, v%update_ghost_nodes (src/solver.f90:291-292) and,
    // two-phase, those of vof, rho, mu, p_o (:299-302)
    for (int f : {FEN_P, FEN_U, FEN_V}) {
        if ((rc = fen_update_ghost_nodes(c, f)) != FEN_OK) return rc;
    }
    if (c->dim == 3 && (rc = fen_update_ghost_nodes(c, FEN_W)) != FEN_OK) return rc;
    if (c->mf) {
        if ((rc = fen_update_ghost_nodes(c, FEN_VOF)) != FEN_OK) return rc;
        try {
            for (double *f : {c->mf->rho, c->mf->mu, c->mf->p_o}) fen::mf_ghost_field(c, f);
            if (hipStreamSynchronize(c->stream) != hipSuccess) throw std::runtime_error("ghost fill failed");
        } catch (const std::exception &e) {
            c->err = e.what();
            return FEN_ERR_HIP;
        }
    }
    return FEN_OK;
}
