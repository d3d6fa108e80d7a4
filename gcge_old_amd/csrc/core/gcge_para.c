/*
 * Solver parameters: defaults, "-gcge_*" command-line flags, setup and the
 * progress/final reports.
 *
 * Defaults and flag names follow reference src/gcge_para.c:25-113 and
 * :128-358; GCGE_PARA_Setup follows :414-457 (block-size defaulting and the
 * block_size <= nev clamp).  Quirks kept on purpose, because they decide
 * which path a given command line selects:
 *   - "-gcge_cg_rate" is read with atoi, so fractional rates become 0
 *     (reference :284-288);
 *   - orth_zero_tol and scbgs_reorth_tol get DBL_EPSILON added in Setup.
 * The flag parser is table driven (one row per flag) instead of a chain of
 * string compares.
 */
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gcge.h"

void GCGE_STATISTIC_PARA_Create(GCGE_STATISTIC_PARA **stat_para)
{
    GCGE_STATISTIC_PARA *s = (GCGE_STATISTIC_PARA *)calloc(1, sizeof(*s));
    s->part_time_total = (GCGE_PART_TIME *)calloc(1, sizeof(GCGE_PART_TIME));
    s->part_time_one_iter = (GCGE_PART_TIME *)calloc(1, sizeof(GCGE_PART_TIME));
    s->ops_time_count_total = (GCGE_OPS_TIME_COUNT *)calloc(1, sizeof(GCGE_OPS_TIME_COUNT));
    s->ops_time_count_linear_solver = (GCGE_OPS_TIME_COUNT *)calloc(1, sizeof(GCGE_OPS_TIME_COUNT));
    *stat_para = s;
}

void GCGE_STATISTIC_PARA_Free(GCGE_STATISTIC_PARA **stat_para)
{
    if (!*stat_para) return;
    free((*stat_para)->part_time_total);
    free((*stat_para)->part_time_one_iter);
    free((*stat_para)->ops_time_count_total);
    free((*stat_para)->ops_time_count_linear_solver);
    free(*stat_para);
    *stat_para = NULL;
}

void GCGE_PARA_Create(GCGE_PARA **para)
{
    GCGE_PARA *p = (GCGE_PARA *)calloc(1, sizeof(GCGE_PARA));
    p->nev = 1;
    p->ev_max_it = 30;
    p->block_size = 0;
    p->if_lobgcg = 0;
    p->num_init_evec = 0;
    p->ev_tol = 1e-8;
    p->eval_type = "sa";
    p->conv_type = "R";
    p->orth_type = "B";
    p->w_orth_type = "multi";
    p->p_orth_type = "multi";
    p->x_orth_type = "bgs";
    p->conv_omega_norm = 0.0;

    p->orth_para = (GCGE_ORTH_PARA *)calloc(1, sizeof(GCGE_ORTH_PARA));
    p->orth_para->orth_zero_tol = 1e-14;
    p->orth_para->reorth_tol = 0.75;
    p->orth_para->scbgs_reorth_tol = 1e-16;
    p->orth_para->max_reorth_time = 3;
    p->orth_para->scbgs_wself_max_reorth_time = 1;
    p->orth_para->criterion_tol = 1e-6;
    p->orth_para->print_orth_zero = 0;
    p->orth_para->x_orth_block_size = 0;
    p->orth_para->w_orth_block_size = 0;
    p->orth_para->max_direct_orth_length = 1;

    p->multi_tol = 0.2;
    p->multi_tol_for_lock = 1e-6;
    p->if_use_cg = 1;
    p->cg_max_it = 20;
    p->cg_rate = 1e-2;
    p->cg_type = 1;
    p->if_use_bcg = 1;
    p->use_bcg_continuous = 0;

    p->num_iter = 0;
    p->num_unlock = 0;
    p->num_conv = 0;
    p->num_conv_continuous = 0;
    p->use_mpi_bcast = 1;

    p->print_cg_error = 0;
    p->print_eval = 0;
    p->print_part_time = 0;
    p->print_final_part_time = 1;
    p->print_para = 1;
    p->print_result = 1;
    p->print_conv = 1;
    p->print_split = 0;

    p->max_res = 0.0;
    p->max_ind = 0;
    p->min_res = 0.0;
    p->min_ind = 0;
    p->sum_res = 0.0;
    p->res = NULL;
    p->dirichlet_boundary = 0;

    GCGE_STATISTIC_PARA_Create(&p->stat_para);

    p->opt_rr_eig_partly = 1;
    p->opt_bcast = 1;
    p->opt_allgatherv = 0;
    p->min_nev_each_proc = 1;
    p->if_shift = 0;
    p->use_shift = 0;
    p->shift = 0.0;
    p->partition_type = "average";

    p->solve_time = 0.0;
    p->use_ops_ext = 1;
    p->full_rr = 0;
    p->svqb_one_pass_cond = 1e4;
    p->svqb_dgks_eta = 0.7071067811865476;
    p->bcg_by_column = 0;
    p->history = NULL;
    p->history_cap = 0;
    p->history_len = 0;
    *para = p;
}

void GCGE_PARA_Free(GCGE_PARA **para)
{
    if (!*para) return;
    GCGE_STATISTIC_PARA_Free(&(*para)->stat_para);
    free((*para)->orth_para);
    free((*para)->res);
    free(*para);
    *para = NULL;
}

/* ---- command line ---------------------------------------------------- */
enum flag_kind { F_INT, F_DBL, F_STR, F_DBL_VIA_ATOI };
typedef struct {
    const char *name;
    enum flag_kind kind;
    int in_orth; /* 1: field lives in para->orth_para */
    size_t offset;
} flag_row;

#define PF(name, kind, field) {name, kind, 0, offsetof(GCGE_PARA, field)}
#define OF(name, kind, field) {name, kind, 1, offsetof(GCGE_ORTH_PARA, field)}

static const flag_row flag_table[] = {
    PF("-gcge_nev", F_INT, nev),
    PF("-gcge_ev_max_it", F_INT, ev_max_it),
    PF("-gcge_block_size", F_INT, block_size),
    OF("-gcge_x_orth_block_size", F_INT, x_orth_block_size),
    OF("-gcge_w_orth_block_size", F_INT, w_orth_block_size),
    OF("-gcge_max_direct_orth_length", F_INT, max_direct_orth_length),
    PF("-gcge_if_lobgcg", F_INT, if_lobgcg),
    PF("-gcge_num_init_evec", F_INT, num_init_evec),
    PF("-gcge_ev_tol", F_DBL, ev_tol),
    PF("-gcge_conv_type", F_STR, conv_type),
    PF("-gcge_eval_type", F_STR, eval_type),
    PF("-gcge_partition_type", F_STR, partition_type),
    PF("-gcge_conv_omega_norm", F_DBL, conv_omega_norm),
    PF("-gcge_orth_type", F_STR, orth_type),
    PF("-gcge_w_orth_type", F_STR, w_orth_type),
    PF("-gcge_p_orth_type", F_STR, p_orth_type),
    PF("-gcge_x_orth_type", F_STR, x_orth_type),
    OF("-gcge_orth_zero_tol", F_DBL, orth_zero_tol),
    OF("-gcge_reorth_tol", F_DBL, reorth_tol),
    OF("-gcge_scbgs_reorth_tol", F_DBL, scbgs_reorth_tol),
    OF("-gcge_max_reorth_time", F_INT, max_reorth_time),
    OF("-gcge_scbgs_wself_max_reorth_time", F_INT, scbgs_wself_max_reorth_time),
    OF("-gcge_print_orth_zero", F_INT, print_orth_zero),
    PF("-gcge_print_split", F_INT, print_split),
    PF("-gcge_multi_tol", F_DBL, multi_tol),
    PF("-gcge_multi_tol_for_lock", F_DBL, multi_tol_for_lock),
    PF("-gcge_if_use_cg", F_INT, if_use_cg),
    PF("-gcge_if_use_bcg", F_INT, if_use_bcg),
    PF("-gcge_use_bcg_continuous", F_INT, use_bcg_continuous),
    PF("-gcge_cg_max_it", F_INT, cg_max_it),
    PF("-gcge_cg_rate", F_DBL_VIA_ATOI, cg_rate),
    PF("-gcge_print_cg_error", F_INT, print_cg_error),
    PF("-gcge_print_eval", F_INT, print_eval),
    PF("-gcge_print_part_time", F_INT, print_part_time),
    PF("-gcge_print_final_part_time", F_INT, print_final_part_time),
    PF("-gcge_print_para", F_INT, print_para),
    PF("-gcge_print_conv", F_INT, print_conv),
    PF("-gcge_print_result", F_INT, print_result),
    PF("-gcge_opt_rr_eig_partly", F_INT, opt_rr_eig_partly),
    PF("-gcge_opt_bcast", F_INT, opt_bcast),
    PF("-gcge_opt_allgatherv", F_INT, opt_allgatherv),
    PF("-gcge_min_nev_each_proc", F_INT, min_nev_each_proc),
    PF("-gcge_use_shift", F_INT, use_shift),
    /* added by this backend */
    PF("-gcge_use_ops_ext", F_INT, use_ops_ext),
    PF("-gcge_full_rr", F_INT, full_rr),
    PF("-gcge_bcg_by_column", F_INT, bcg_by_column),
    PF("-gcge_svqb_one_pass_cond", F_DBL, svqb_one_pass_cond),
    PF("-gcge_svqb_dgks_eta", F_DBL, svqb_dgks_eta),
    /* a GCGE_PARA field the reference sets only in code (example/csr_solver.c:51) */
    PF("-gcge_dirichlet_boundary", F_INT, dirichlet_boundary),
};

static void print_usage(const char *prog)
{
    GCGE_Printf("\nUsage: %s [<options>]\n\n", prog);
    GCGE_Printf("  -gcge_nev <i>               number of eigenpairs (default 1)\n");
    GCGE_Printf("  -gcge_ev_max_it <i>         maximum GCG iterations (default 30)\n");
    GCGE_Printf("  -gcge_block_size <i>        unconverged pairs per batch (default max(min(nev,10),nev/5))\n");
    GCGE_Printf("  -gcge_ev_tol <d>            convergence tolerance (default 1e-8)\n");
    GCGE_Printf("  -gcge_conv_type <c>         R | A | O | RB (default R)\n");
    GCGE_Printf("  -gcge_orth_type <c>         B or A inner product (default B)\n");
    GCGE_Printf("  -gcge_x_orth_type <c>       bgs | multi | cbgs | scbgs (default bgs)\n");
    GCGE_Printf("  -gcge_w_orth_type <c>       multi | bgs | cbgs | scbgs | gs (default multi)\n");
    GCGE_Printf("  -gcge_p_orth_type <c>       multi | bgs | gs (default multi)\n");
    GCGE_Printf("  -gcge_cg_max_it <i>         inner CG iterations (default 20)\n");
    GCGE_Printf("  -gcge_cg_rate <i>           inner CG reduction (parsed as integer; default 1e-2)\n");
    GCGE_Printf("  -gcge_use_ops_ext <i>       use backend block extensions (default 1)\n");
    GCGE_Printf("  -gcge_full_rr <i>           keep converged vectors in the Rayleigh-Ritz step (default 0)\n");
    GCGE_Printf("  -gcge_bcg_by_column <i>     block CG scalars kept per column (default 0)\n");
    GCGE_Printf("  -gcge_svqb_one_pass_cond <d> one SVQB pass below this Gram condition (default 1e4; 0: two)\n");
    GCGE_Printf("  -gcge_svqb_dgks_eta <d>    svqb: W as a correction, second projection pass below this\n"
                "                             kept B-norm fraction (default 0.7071; 0: reference W, all passes)\n");
    GCGE_Printf("  -gcge_dirichlet_boundary <i> zero the Dirichlet rows of the initial vectors (default 0)\n");
    GCGE_Printf("  (all other reference -gcge_* flags are accepted with the same meaning)\n\n");
}

GCGE_INT GCGE_PARA_SetFromCommandLine(GCGE_PARA *para, GCGE_INT argc, char **argv)
{
    int i = 0, help = 0;
    const int nflags = (int)(sizeof(flag_table) / sizeof(flag_table[0]));
    while (i < argc) {
        if (strcmp(argv[i], "-gcge_help") == 0) {
            help = 1;
            break;
        }
        int matched = 0;
        for (int f = 0; f < nflags; ++f) {
            if (strcmp(argv[i], flag_table[f].name) != 0) continue;
            matched = 1;
            if (i + 1 >= argc) { i = argc; break; }
            char *base = flag_table[f].in_orth ? (char *)para->orth_para : (char *)para;
            void *field = base + flag_table[f].offset;
            const char *val = argv[i + 1];
            switch (flag_table[f].kind) {
            case F_INT: *(GCGE_INT *)field = atoi(val); break;
            case F_DBL: *(GCGE_DOUBLE *)field = atof(val); break;
            case F_DBL_VIA_ATOI: *(GCGE_DOUBLE *)field = (GCGE_DOUBLE)atoi(val); break;
            case F_STR: *(char **)field = (char *)val; break;
            }
            i += 2;
            break;
        }
        if (!matched) ++i;
    }
    if (help) {
        print_usage(argc > 0 ? argv[0] : "gcge");
        exit(0);
    }
    return 0;
}

void GCGE_PARA_SetNumEigen(GCGE_PARA *para, GCGE_INT nev) { para->nev = nev; }

void GCGE_PARA_Setup(GCGE_PARA *para)
{
    GCGE_INT nev = para->nev;
    if (para->block_size == 0) {
        GCGE_INT bs = nev < 10 ? nev : 10;
        para->block_size = (nev / 5 > bs) ? nev / 5 : bs;
    } else if (para->block_size > nev) {
        para->block_size = nev;
    }
    if (para->orth_para->x_orth_block_size == 0)
        para->orth_para->x_orth_block_size = nev / 5;
    if (para->orth_para->w_orth_block_size == 0)
        para->orth_para->w_orth_block_size = (para->block_size / 5 > 1) ? para->block_size / 5 : 1;
    para->orth_para->orth_zero_tol += DBL_EPSILON;
    para->orth_para->scbgs_reorth_tol += DBL_EPSILON;
    free(para->res);
    para->res = (GCGE_DOUBLE *)calloc((size_t)nev, sizeof(GCGE_DOUBLE));
    for (GCGE_INT i = 0; i < nev; ++i) para->res[i] = 1.0;
}

/* ---- reports (same line formats as reference src/gcge_para.c:461-754) ---- */
static const char *residual_name(const char *conv_type)
{
    if (strcmp(conv_type, "R") == 0) return "relative_res";
    if (strcmp(conv_type, "O") == 0) return "omega_res";
    if (strcmp(conv_type, "RB") == 0) return "b_relative_res";
    return "abosolute_res";
}

static GCGE_DOUBLE sum_part(const GCGE_PART_TIME *t)
{
    return t->x_axpy_time + t->p_axpy_time + t->x_orth_time + t->p_orth_time + t->w_line_time +
           t->w_orth_time + t->rr_mat_time + t->rr_eigen_time + t->conv_time;
}

void GCGE_PrintIterationInfo(GCGE_DOUBLE *eval, GCGE_PARA *para)
{
    GCGE_INT it = para->num_iter;
    if (para->print_conv == 1)
        GCGE_Printf("num_conv_continuous(%3d) = %3d; max_res(%3d) = %e; min_res(%3d) = %e;\n", it,
                    para->num_conv_continuous, it, para->max_res, it, para->min_res);
    if (para->print_eval == 1) {
        const char *name = residual_name(para->conv_type);
        for (GCGE_INT i = 0; i < para->nev; ++i) {
            GCGE_Printf("eval(%d,%3d) = %20.15e ; %s(%d,%3d) =  %20.15e ;", it, i + 1, eval[i], name,
                        it, i + 1, para->res[i]);
            if (para->res[i] < para->ev_tol) GCGE_Printf(" (converged)\n");
            else if (i == (GCGE_INT)para->max_ind) GCGE_Printf(" (max)\n");
            else if (i == (GCGE_INT)para->min_ind) GCGE_Printf(" (min)\n");
            else GCGE_Printf("\n");
        }
    }
    if (para->print_part_time == 1) {
        GCGE_PART_TIME *t = para->stat_para->part_time_one_iter;
        GCGE_Printf("\niteration: %d, time: %12.4lf\n", it, sum_part(t));
        GCGE_Printf("  x_orth %.4f x_axpy %.4f p_orth %.4f p_axpy %.4f check_conv %.4f\n",
                    t->x_orth_time, t->x_axpy_time, t->p_orth_time, t->p_axpy_time, t->conv_time);
        GCGE_Printf("  w_lsol %.4f w_orth %.4f rr_eig %.4f rr_mat %.4f\n", t->w_line_time,
                    t->w_orth_time, t->rr_eigen_time, t->rr_mat_time);
    }
    fflush(stdout);
}

void GCGE_PrintFinalInfo(GCGE_DOUBLE *eval, GCGE_PARA *para)
{
    GCGE_PART_TIME *t = para->stat_para->part_time_total;
    GCGE_DOUBLE total = sum_part(t);
    if (para->print_result) {
        GCGE_Printf("\nFinal eigenvalues and residual information:\n\n            eigenvalue           %s\n\n",
                    residual_name(para->conv_type));
        for (GCGE_INT i = 0; i < para->nev; ++i) {
            GCGE_Printf("%5d  %20.15e  %20.15e ", i + 1, eval[i], para->res[i]);
            GCGE_Printf(para->res[i] < para->ev_tol ? " (converged)\n" : "\n");
        }
    }
    if (para->print_conv == 1) {
        GCGE_Printf("\nTotal gcg iteration: %d, total time: %.4lf. ", para->num_iter, total);
        if (para->num_conv_continuous == para->nev)
            GCGE_Printf("All %d eigenpairs converged! ", para->nev);
        else
            GCGE_Printf("%d converged, %d unconverged. ", para->num_conv_continuous,
                        para->nev - para->num_conv_continuous);
    }
    if (para->print_final_part_time == 1) {
        GCGE_Printf("Use 1 process. block_size: %d\n", para->block_size);
        GCGE_Printf("Section Name        x_orth       x_axpy       p_orth       p_axpy   check_conv\n");
        GCGE_Printf("Section Time  %12.4lf %12.4lf %12.4lf %12.4lf %12.4lf\n", t->x_orth_time,
                    t->x_axpy_time, t->p_orth_time, t->p_axpy_time, t->conv_time);
        GCGE_Printf("Section Name        w_lsol       w_orth       rr_eig       rr_mat\n");
        GCGE_Printf("Section Time  %12.4lf %12.4lf %12.4lf %12.4lf\n", t->w_line_time,
                    t->w_orth_time, t->rr_eigen_time, t->rr_mat_time);
        if (total >= 1e-10)
            GCGE_Printf("Section Ratio %11.2lf%% %11.2lf%% %11.2lf%% %11.2lf%%\n\n",
                        100 * t->w_line_time / total, 100 * t->w_orth_time / total,
                        100 * t->rr_eigen_time / total, 100 * t->rr_mat_time / total);
    }
    fflush(stdout);
}

void GCGE_PrintParaInfo(GCGE_PARA *para)
{
    if (para->print_para != 1) return;
    GCGE_ORTH_PARA *o = para->orth_para;
    GCGE_Printf("\nGCGE_PARA:\n\n");
    GCGE_Printf("  nev                        : %8d\n", para->nev);
    GCGE_Printf("  ev_max_it                  : %8d\n", para->ev_max_it);
    GCGE_Printf("  block_size                 : %8d\n", para->block_size);
    GCGE_Printf("  num_init_evec              : %8d\n", para->num_init_evec);
    GCGE_Printf("  max_reorth_time            : %8d\n", o->max_reorth_time);
    GCGE_Printf("  if_use_bcg                 : %8d\n", para->if_use_bcg);
    GCGE_Printf("  use_bcg_continuous         : %8d\n", para->use_bcg_continuous);
    GCGE_Printf("  cg_max_it                  : %8d\n", para->cg_max_it);
    GCGE_Printf("  orth_type                  : %8s\n", para->orth_type);
    GCGE_Printf("  w_orth_type                : %8s\n", para->w_orth_type);
    GCGE_Printf("  p_orth_type                : %8s\n", para->p_orth_type);
    GCGE_Printf("  x_orth_type                : %8s\n", para->x_orth_type);
    GCGE_Printf("  max_direct_orth_length     : %8d\n", o->max_direct_orth_length);
    GCGE_Printf("  conv_type                  : %8s\n", para->conv_type);
    GCGE_Printf("  eval_type                  : %8s\n", para->eval_type);
    GCGE_Printf("  ev_tol                     : %3.2e\n", para->ev_tol);
    GCGE_Printf("  orth_zero_tol              : %3.2e\n", o->orth_zero_tol);
    GCGE_Printf("  reorth_tol                 : %3.2e\n", o->reorth_tol);
    GCGE_Printf("  scbgs_reorth_tol           : %3.2e\n", o->scbgs_reorth_tol);
    GCGE_Printf("  cg_rate                    : %3.2e\n", para->cg_rate);
    GCGE_Printf("  multi_tol_for_lock         : %3.2e\n", para->multi_tol_for_lock);
    GCGE_Printf("  use_shift                  : %8d\n", para->use_shift);
    GCGE_Printf("  opt_rr_eig_partly          : %8d\n", para->opt_rr_eig_partly);
    GCGE_Printf("  use_ops_ext                : %8d\n", para->use_ops_ext);
    GCGE_Printf("  full_rr                    : %8d\n", para->full_rr);
    GCGE_Printf("  bcg_by_column              : %8d\n", para->bcg_by_column);
    GCGE_Printf("  svqb_one_pass_cond         : %3.2e\n", para->svqb_one_pass_cond);
    GCGE_Printf("  svqb_dgks_eta              : %3.2e\n\n", para->svqb_dgks_eta);
}
