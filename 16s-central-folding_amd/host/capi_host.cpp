// C entry points of the host side: system construction and error reporting.
#include <cstdlib>
#include <new>
#include <string>

#include "../../include/tis_gpu.h"
#include "glibc_rand.h"
#include "host_util.h"

namespace tis_host {
tis_system *build_16scd(const std::string &dir, const tis_build_args &args, long long *draws);
tis_system *build_synthetic(const std::string &dir, const tis_synthetic_args &args);
tis_system *build_chain(const tis_chain_files &f, const tis_build_args &args, long long *draws);

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
}  // namespace tis_host

extern "C" const char *tis_last_error(void) { return tis_host::g_error.c_str(); }

extern "C" int tis_build_16scd(const char *data_dir, const tis_build_args *args, tis_system **out,
                               long long *rand_draws_out) {
    if (!args || !out) { tis_host::set_error("tis_build_16scd: null argument"); return TIS_EINVAL; }
    *out = nullptr;
    if (!(args->T > 0) || args->c_mg < 0 || args->c_k < 0 || !(args->box >= 0)) {
        tis_host::set_error("tis_build_16scd: T must be > 0 and concentrations >= 0");
        return TIS_EINVAL;
    }
    try {
        *out = tis_host::build_16scd(data_dir ? data_dir : ".", *args, rand_draws_out);
        return TIS_OK;
    } catch (const tis_host::IoError &e) {
        tis_host::set_error(e.what());
        return TIS_EIO;
    } catch (const std::bad_alloc &) {
        tis_host::set_error("tis_build_16scd: out of host memory");
        return TIS_ENOMEM;
    } catch (const std::exception &e) {
        tis_host::set_error(e.what());
        return TIS_EIO;
    }
}

extern "C" int tis_build_chain(const tis_chain_files *files, const tis_build_args *args, tis_system **out,
                               long long *rand_draws_out) {
    if (!files || !args || !out || !files->pdb || !files->bonds || !files->stacks || !files->maxikey ||
        (files->n_double_face && !files->double_face_res) || files->n_double_face < 0) {
        tis_host::set_error("tis_build_chain: null or inconsistent argument");
        return TIS_EINVAL;
    }
    *out = nullptr;
    if (!(args->T > 0) || args->c_mg < 0 || args->c_k < 0 || !(args->box >= 0)) {
        tis_host::set_error("tis_build_chain: T must be > 0 and concentrations >= 0");
        return TIS_EINVAL;
    }
    try {
        *out = tis_host::build_chain(*files, *args, rand_draws_out);
        return TIS_OK;
    } catch (const tis_host::IoError &e) {
        tis_host::set_error(e.what());
        return TIS_EIO;
    } catch (const std::bad_alloc &) {
        tis_host::set_error("tis_build_chain: out of host memory");
        return TIS_ENOMEM;
    } catch (const std::exception &e) {
        tis_host::set_error(e.what());
        return TIS_EIO;
    }
}

extern "C" int tis_build_synthetic(const char *data_dir, const tis_synthetic_args *args, tis_system **out) {
    if (!args || !out) { tis_host::set_error("tis_build_synthetic: null argument"); return TIS_EINVAL; }
    *out = nullptr;
    const tis_build_args &b = args->build;
    if (args->n_nt < 8 || !(args->radius > 0) || !(args->stack_frac >= 0 && args->stack_frac <= 1) ||
        args->n_hb < 0 || !(b.T > 0) || b.c_mg < 0 || b.c_k < 0 || !(b.box >= 5 * args->radius)) {
        tis_host::set_error("tis_build_synthetic: need n_nt >= 8, radius > 0, 0 <= stack_frac <= 1, "
                            "n_hb >= 0, T > 0, concentrations >= 0 and box >= 5 * radius");
        return TIS_EINVAL;
    }
    try {
        *out = tis_host::build_synthetic(data_dir ? data_dir : ".", *args);
        return TIS_OK;
    } catch (const tis_host::IoError &e) {
        tis_host::set_error(e.what());
        return TIS_EIO;
    } catch (const std::bad_alloc &) {
        tis_host::set_error("tis_build_synthetic: out of host memory");
        return TIS_ENOMEM;
    } catch (const std::exception &e) {
        tis_host::set_error(e.what());
        return TIS_EIO;
    }
}

extern "C" void tis_system_free(tis_system *s) {
    if (!s) return;
    void *ptrs[] = {s->type, s->cls, s->pos, s->pos_old, s->vel, s->unit_bead0, s->unit_bead1,
                    s->unit_kind, s->bond_ij, s->bond_kr, s->angle_ijk, s->angle_kt, s->st3_beads,
                    s->st3_geo, s->st3_E, s->st3_face, s->st3_disable, s->st2_beads, s->st2_r0,
                    s->st2_E, s->st_excess, s->hb_beads, s->hb_geo, s->hb_E, s->hb_dode,
                    s->hb_site_off, s->hb_site_idx, s->site_hb_off, s->site_hb_idx, s->site_valence};
    for (void *p : ptrs) std::free(p);
    std::free(s);
}

extern "C" int tis_glibc_rand_fill(unsigned seed, long long skip, int32_t *out, long long n) {
    if ((n && !out) || n < 0 || skip < 0) { tis_host::set_error("tis_glibc_rand_fill: bad arguments"); return TIS_EINVAL; }
    tis_host::GlibcRand g(seed);
    for (long long q = 0; q < skip; ++q) g.next();
    for (long long q = 0; q < n; ++q) out[q] = g.next();
    return TIS_OK;
}
