// input.cpp -- the reference's input.dat reader and the parameter setters
// that copy its values into the solver structures, plus the raw-pointer
// wrapper that solves with parameters from ./input.dat.
//
// File format (src/utilities/input.c:42-830): whitespace-separated tokens;
// a token starting with '[', '%' or '|' comments out the rest of its line;
// otherwise the token is a keyword that must be followed by the token "=" and
// a value read like scanf's %lld (integers), %lf (reals) or %s (names); the
// rest of the line is then skipped.  An unknown keyword prints a warning and
// skips its line; a malformed entry ends the reading silently with the values
// read so far (the reference's behaviour).
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "hazmath_mi355x.h"

namespace hz {
int fatal_status(int code);
}

namespace {

enum class Field { Int, Short, Real, Text, AmgType, Cycle, Smoother, OnOff };

struct Key {
    const char *name;
    Field kind;
    size_t offset;
};

#define HZ_KEY(name, kind) {#name, Field::kind, offsetof(input_param, name)}

const Key kKeys[] = {
    HZ_KEY(print_level, Short),
    HZ_KEY(output_dir, Text),
    HZ_KEY(read_mesh_from_file, Int),
    HZ_KEY(gridfile, Text),
    HZ_KEY(spatial_dim, Int),
    HZ_KEY(refinement_type, Int),
    HZ_KEY(refinement_levels, Int),
    HZ_KEY(boundary_codes, Int),
    HZ_KEY(nquad, Int),
    HZ_KEY(FE_type, Int),
    HZ_KEY(Mass_lump, Int),
    HZ_KEY(time_start, Real),
    HZ_KEY(time_step_type, Int),
    HZ_KEY(time_steps, Int),
    HZ_KEY(time_step_size, Real),
    HZ_KEY(rhs_time_dep, Int),
    HZ_KEY(nonlinear_itsolver_type, Int),
    HZ_KEY(nonlinear_itsolver_maxit, Int),
    HZ_KEY(nonlinear_itsolver_tol, Real),
    HZ_KEY(nonlinear_itsolver_toltype, Int),
    HZ_KEY(fas_presmoothers, Int),
    HZ_KEY(fas_postsmoothers, Int),
    // fas_smooth_tol is an INT field the reference fills from a %lf read
    {"fas_smooth_tol", Field::Real, (size_t)-1},
    HZ_KEY(linear_itsolver_type, Int),
    HZ_KEY(linear_itsolver_maxit, Int),
    HZ_KEY(linear_itsolver_tol, Real),
    HZ_KEY(linear_stop_type, Short),
    HZ_KEY(linear_restart, Int),
    HZ_KEY(linear_precond_type, Int),
    HZ_KEY(AMG_tol, Real),
    HZ_KEY(AMG_maxit, Int),
    HZ_KEY(AMG_type, AmgType),
    HZ_KEY(AMG_levels, Short),
    HZ_KEY(AMG_cycle_type, Cycle),
    HZ_KEY(AMG_amli_degree, Short),
    HZ_KEY(AMG_nl_amli_krylov_type, Short),
    HZ_KEY(AMG_smoother, Smoother),
    HZ_KEY(AMG_relaxation, Real),
    HZ_KEY(AMG_Schwarz_levels, Int),
    HZ_KEY(AMG_presmooth_iter, Short),
    HZ_KEY(AMG_postsmooth_iter, Short),
    HZ_KEY(AMG_coarse_dof, Int),
    HZ_KEY(AMG_coarse_solver, Short),
    HZ_KEY(AMG_coarse_scaling, OnOff),
    HZ_KEY(AMG_fpwr, Real),
    HZ_KEY(AMG_aggregation_type, Short),
    HZ_KEY(AMG_strong_coupled, Real),
    HZ_KEY(AMG_max_aggregation, Int),
    HZ_KEY(AMG_tentative_smooth, Real),
    HZ_KEY(AMG_smooth_filter, OnOff),
    HZ_KEY(Schwarz_mmsize, Int),
    HZ_KEY(Schwarz_maxlvl, Int),
    HZ_KEY(Schwarz_type, Int),
    HZ_KEY(Schwarz_blksolver, Int),
    HZ_KEY(HX_smooth_iter, Short),
    HZ_KEY(BSR_alpha, Real),
    HZ_KEY(BSR_omega, Real),
};

#undef HZ_KEY

struct Name {
    const char *upper, *lower;
    int value;
};

// include/macro.h values
const Name kAmgTypes[] = {{"UA", "ua", 1}, {"SA", "sa", 2}, {"C", "c", 3}, {"MUA", "mua", 10}, {"MSA", "msa", 11}};
const Name kCycles[] = {{"V", "v", 1}, {"W", "w", 2}, {"A", "a", 3}, {"NA", "na", 4}, {"ADD", "add", 5}};
const Name kSmoothers[] = {{"JACOBI", "jacobi", 1},        {"GS", "gs", 2},
                           {"SGS", "sgs", 3},              {"CG", "cg", 4},
                           {"SOR", "sor", 5},              {"SSOR", "ssor", 6},
                           {"GSOR", "gsor", 7},            {"SGSOR", "sgsor", 8},
                           {"POLY", "poly", 9},            {"L1DIAG", "l1diag", 10},
                           {"FJACOBI", "fjacobi", 11},     {"FGS", "fgs", 12},
                           {"FSGS", "fsgs", 13},           {"JACOBI_GS", "jacobi_gs", 31},
                           {"JACOBI_SGS", "jacobi_sgs", 32}, {"METRIC_ADD", "metric_add", 40},
                           {"METRIC_MUL", "metric_mul", 41}};

template <size_t N>
bool lookup(const Name (&table)[N], const char *s, int *out) {
    for (const Name &e : table)
        if (std::strcmp(s, e.upper) == 0 || std::strcmp(s, e.lower) == 0) {
            *out = e.value;
            return true;
        }
    return false;
}

// ON in any letter case, OFF in any letter case
bool on_off(const char *s, int *out) {
    char t[8];
    size_t k = 0;
    for (; s[k] && k < sizeof(t) - 1; ++k) t[k] = (char)((s[k] >= 'A' && s[k] <= 'Z') ? s[k] - 'A' + 'a' : s[k]);
    t[k] = 0;
    if (s[k]) return false;
    if (std::strcmp(t, "on") == 0) { *out = 1; return true; }
    if (std::strcmp(t, "off") == 0) { *out = 0; return true; }
    return false;
}

void skip_line(FILE *fp) {
    int c;
    while ((c = std::fgetc(fp)) != EOF && c != '\n') {
    }
}

template <class T>
void put(input_param *p, size_t off, T v) {
    std::memcpy(reinterpret_cast<char *>(p) + off, &v, sizeof(T));
}

// Reads "= value" for key k into p; false on a malformed entry.
bool read_entry(FILE *fp, const Key &k, input_param *p) {
    char tok[512];
    if (std::fscanf(fp, "%511s", tok) != 1 || std::strcmp(tok, "=") != 0) return false;
    switch (k.kind) {
        case Field::Int:
        case Field::Short: {
            long long v = 0;
            if (std::fscanf(fp, "%lld", &v) != 1) return false;
            const INT iv = (INT)v;
            if (k.kind == Field::Int) put<INT>(p, k.offset, iv);
            else put<SHORT>(p, k.offset, (SHORT)iv);
            break;
        }
        case Field::Real: {
            double v = 0.0;
            if (std::fscanf(fp, "%lf", &v) != 1) return false;
            if (k.offset == (size_t)-1) put<INT>(p, offsetof(input_param, fas_smooth_tol), (INT)v);
            else put<REAL>(p, k.offset, v);
            break;
        }
        case Field::Text: {
            if (std::fscanf(fp, "%511s", tok) != 1) return false;
            char *dst = reinterpret_cast<char *>(p) + k.offset;
            std::strncpy(dst, tok, 128);
            break;
        }
        default: {
            if (std::fscanf(fp, "%511s", tok) != 1) return false;
            int v = 0;
            bool ok = false;
            if (k.kind == Field::AmgType) ok = lookup(kAmgTypes, tok, &v);
            else if (k.kind == Field::Cycle) ok = lookup(kCycles, tok, &v);
            else if (k.kind == Field::Smoother) ok = lookup(kSmoothers, tok, &v);
            else ok = on_off(tok, &v);
            if (!ok) return false;
            put<SHORT>(p, k.offset, (SHORT)v);
            break;
        }
    }
    skip_line(fp);
    return true;
}

}  // namespace

extern "C" {

// parameters.c:39-141
void param_input_init(input_param *p) {
    std::memset(p, 0, sizeof(*p));
    p->print_level = 2;
    std::strcpy(p->inifile, "./input.dat");
    std::strcpy(p->gridfile, "../grids/2D/unitSQ_hp125.dat");
    std::strcpy(p->output_dir, "./output/");
    p->read_mesh_from_file = 0;
    p->spatial_dim = 2;
    p->refinement_type = 11;
    p->refinement_levels = 2;
    p->boundary_codes = 1;
    p->nquad = 2;
    p->FE_type = 1;
    p->Mass_lump = 0;
    p->time_start = 0.0;
    p->time_step_type = 0;
    p->time_steps = 0;
    p->time_step_size = 0.01;
    p->rhs_time_dep = 1;
    p->nonlinear_itsolver_type = 0;
    p->nonlinear_itsolver_maxit = 0;
    p->nonlinear_itsolver_tol = 1e-6;
    p->nonlinear_itsolver_toltype = 0;
    p->linear_itsolver_type = 3;   // SOLVER_VGMRES
    p->linear_precond_type = 0;    // PREC_NULL
    p->linear_stop_type = HZ_STOP_REL_RES;
    p->linear_itsolver_tol = 1e-6;
    p->linear_itsolver_maxit = 500;
    p->linear_restart = 25;
    p->AMG_type = HZ_UA_AMG;
    p->AMG_levels = 10;
    p->AMG_cycle_type = HZ_W_CYCLE;
    p->AMG_smoother = HZ_SMOOTHER_GS;
    p->AMG_presmooth_iter = 1;
    p->AMG_postsmooth_iter = 1;
    p->AMG_polynomial_degree = 2;
    p->AMG_relaxation = 1.2;
    p->AMG_coarse_dof = 200;
    p->AMG_coarse_solver = HZ_SOLVER_DEFAULT;
    p->AMG_tol = 1e-6;
    p->AMG_maxit = 1;
    p->AMG_Schwarz_levels = 0;
    p->AMG_coarse_scaling = 0;
    p->AMG_amli_degree = 1;
    p->AMG_nl_amli_krylov_type = 2;
    p->AMG_fpwr = 1.0;
    p->AMG_aggregation_type = HZ_HEC;
    p->AMG_strong_coupled = 0.04;
    p->AMG_max_aggregation = 20;
    p->AMG_tentative_smooth = 0.67;
    p->AMG_smooth_filter = 1;
    p->Schwarz_mmsize = 200;
    p->Schwarz_maxlvl = 2;
    p->Schwarz_type = 1;
    p->Schwarz_blksolver = HZ_SOLVER_DEFAULT;
    p->HX_smooth_iter = 1;
    p->BSR_alpha = 1.0;
    p->BSR_omega = 1.0;
}

// input.c:42-830
void param_input(const char *filenm, input_param *p) {
    param_input_init(p);
    if (filenm == nullptr) return;
    std::strncpy(p->inifile, filenm, sizeof(p->inifile) - 1);
    FILE *fp = std::fopen(filenm, "r");
    if (!fp) {
        std::printf("### ERROR: Cannot open file %s!\n", filenm);
        hz::fatal_status(HZ_ERROR_OPEN_FILE);
        return;
    }
    char tok[512];
    while (std::fscanf(fp, "%511s", tok) == 1) {
        if (tok[0] == '[' || tok[0] == '%' || tok[0] == '|') {
            skip_line(fp);
            continue;
        }
        const Key *k = nullptr;
        for (const Key &e : kKeys)
            if (std::strcmp(tok, e.name) == 0) {
                k = &e;
                break;
            }
        if (!k) {
            std::printf("### HAZMATH WARNING: Unknown input keyword %s!\n", tok);
            skip_line(fp);
            continue;
        }
        if (!read_entry(fp, *k, p)) break;
    }
    std::fclose(fp);
    std::fflush(stdout);
}

// parameters.c:263-283
void param_linear_solver_set(linear_itsolver_param *it, input_param *p) {
    it->linear_print_level = p->print_level;
    it->linear_itsolver_type = (SHORT)p->linear_itsolver_type;
    it->linear_stop_type = p->linear_stop_type;
    it->linear_restart = p->linear_restart;
    it->linear_precond_type = (SHORT)p->linear_precond_type;
    if (it->linear_itsolver_type == HZ_SOLVER_AMG) {
        it->linear_tol = p->AMG_tol;
        it->linear_maxit = p->AMG_maxit;
    } else {
        it->linear_tol = p->linear_itsolver_tol;
        it->linear_maxit = p->linear_itsolver_maxit;
    }
    it->HX_smooth_iter = p->HX_smooth_iter;
}

// parameters.c:296-348
void param_amg_set(AMG_param *a, input_param *p) {
    a->AMG_type = p->AMG_type;
    a->print_level = p->print_level;
    if (p->linear_itsolver_type == HZ_SOLVER_AMG) {
        a->maxit = p->linear_itsolver_maxit;
        a->tol = p->linear_itsolver_tol;
    } else {
        a->maxit = p->AMG_maxit;
        a->tol = p->AMG_tol;
    }
    a->max_levels = p->AMG_levels;
    a->cycle_type = p->AMG_cycle_type;
    a->smoother = p->AMG_smoother;
    a->relaxation = p->AMG_relaxation;
    a->polynomial_degree = p->AMG_polynomial_degree;
    a->presmooth_iter = p->AMG_presmooth_iter;
    a->postsmooth_iter = p->AMG_postsmooth_iter;
    a->coarse_solver = p->AMG_coarse_solver;
    a->coarse_dof = p->AMG_coarse_dof;
    a->coarse_scaling = p->AMG_coarse_scaling;
    a->amli_degree = p->AMG_amli_degree;
    a->amli_coef = nullptr;
    a->nl_amli_krylov_type = p->AMG_nl_amli_krylov_type;
    a->fpwr = p->AMG_fpwr;
    a->aggregation_type = p->AMG_aggregation_type;
    a->strong_coupled = p->AMG_strong_coupled;
    a->max_aggregation = p->AMG_max_aggregation;
    a->tentative_smooth = p->AMG_tentative_smooth;
    a->smooth_filter = p->AMG_smooth_filter;
    a->Schwarz_levels = p->AMG_Schwarz_levels;
    a->Schwarz_mmsize = p->Schwarz_mmsize;
    a->Schwarz_maxlvl = p->Schwarz_maxlvl;
    a->Schwarz_type = p->Schwarz_type;
    a->Schwarz_blksolver = p->Schwarz_blksolver;
    a->Schwarz_on_blk = nullptr;
    a->HAZDIR = nullptr;
    a->Schwarz_patch_type = nullptr;
    a->damping_param = 0.0;
    a->BSR_alpha = p->BSR_alpha;
    a->BSR_omega = p->BSR_omega;
}

// wrapper.c:36-79: parameters from ./input.dat, then the Krylov-AMG solve on
// the caller's arrays (no copy of A; u is overwritten).
void python_wrapper_krylov_amg(INT *n, INT *nnz, INT *ia, INT *ja, REAL *a, REAL *b, REAL *u, REAL *tol,
                               INT *maxit, INT *print_lvl) {
    input_param inparam;
    param_input("./input.dat", &inparam);
    linear_itsolver_param itparam;
    AMG_param amgparam;
    param_linear_solver_init(&itparam);
    param_linear_solver_set(&itparam, &inparam);
    param_amg_init(&amgparam);
    param_amg_set(&amgparam, &inparam);
    amgparam.print_level = (SHORT)*print_lvl;
    itparam.linear_tol = *tol;
    itparam.linear_print_level = (SHORT)*print_lvl;
    itparam.linear_maxit = *maxit;
    dCSRmat mat;
    mat.row = *n;
    mat.col = *n;
    mat.nnz = *nnz;
    mat.IA = ia;
    mat.JA = ja;
    mat.val = a;
    dvector rhs, sol;
    rhs.row = *n;
    rhs.val = b;
    sol.row = *n;
    sol.val = u;
    linear_solver_dcsr_krylov_amg(&mat, &rhs, &sol, &itparam, &amgparam);
}

}  // extern "C"
