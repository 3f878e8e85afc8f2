/*
 * Test infrastructure only -- the oracle. Only tests/, the smoke check and the
 * benchmark's cpu_baseline leg may load it; the product never links it.
 *
 * Precision-generic body of the CPU restatement of the reference iterations.
 * Included twice by emsa_oracle.c: once with REAL=float (the FP32 restatement
 * of emicp_cpu, the timed CPU baseline) and once with REAL=double (the FP64
 * oracle of the identical iteration).
 *
 * Required macros: REAL, SFX (name suffix), EXPR (exp), SQRTR (sqrt),
 * FABSR (fabs), FIND_RT (Procrustes for this precision).
 */

/* ------------------------------------------------------------------------- */
/* EM-ICP, one sigma step.                                                     */
/* Follows src/emicp_cpu.cpp:60-126 (updateA, normalizeRowsOfA) and           */
/* src/emicp_cpu.cpp:342-505 (row sums, outlier term, lambda, X', centroids,   */
/* centring, S, findRTfromS); identical maths in src/emicp.cu:443-668.        */
/* Rows r run over Y (M), columns c over X (N), as in emicp.cu:337-355.       */
/* The N x M matrix is streamed one row at a time: every element is computed  */
/* with the same operations as the reference, only its storage differs.       */
/* ------------------------------------------------------------------------- */
static void CAT(emicp_rows_, SFX)(const REAL *X, int N, const REAL *Y, int M,
                                   const REAL *R, const REAL *t, float sigma_p2,
                                   REAL outlier, REAL *rowC, REAL *rowLambda,
                                   REAL *Xp /* 3*M SoA */)
{
    const REAL *Xx = X, *Xy = X + N, *Xz = X + 2 * (size_t)N;
    const REAL *Yx = Y, *Yy = Y + M, *Yz = Y + 2 * (size_t)M;
    const REAL sig = (REAL)sigma_p2;
    const REAL floorA = (REAL)(1.0f / (float)N); /* emicp_cpu.cpp:125 */
#pragma omp parallel
    {
        REAL *A = (REAL *)malloc(sizeof(REAL) * (size_t)(N > 0 ? N : 1));
#pragma omp for schedule(static)
        for (int r = 0; r < M; r++) {
            /* emicp_cpu.cpp:92-104: same expression per pair, hoisted */
            REAL px = R[0] * Yx[r] + R[1] * Yy[r] + R[2] * Yz[r] + t[0];
            REAL py = R[3] * Yx[r] + R[4] * Yy[r] + R[5] * Yz[r] + t[1];
            REAL pz = R[6] * Yx[r] + R[7] * Yy[r] + R[8] * Yz[r] + t[2];
            for (int c = 0; c < N; c++) {
                REAL tx = Xx[c] - px, ty = Xy[c] - py, tz = Xz[c] - pz;
                tx *= tx;
                ty *= ty;
                tz *= tz;
                tx += ty;
                tx += tz;
                tx /= sig;
                A[c] = EXPR(-tx);
            }
            /* emicp_cpu.cpp:342-358: C = A*1 + exp(-d02/sigma) */
            REAL C = 0;
            for (int c = 0; c < N; c++) C += A[c];
            C += outlier;
            /* emicp_cpu.cpp:121-125: floor compared in float (10e-7f) */
            if ((float)C > 10e-7f)
                for (int c = 0; c < N; c++) A[c] = SQRTR(A[c] / C);
            else
                for (int c = 0; c < N; c++) A[c] = floorA;
            /* emicp_cpu.cpp:380-386: lambda = A'*1 */
            REAL lam = 0;
            for (int c = 0; c < N; c++) lam += A[c];
            /* emicp_cpu.cpp:412-420: X' = A'X, then X' ./ lambda */
            REAL sx = 0, sy = 0, sz = 0;
            for (int c = 0; c < N; c++) {
                sx += A[c] * Xx[c];
                sy += A[c] * Xy[c];
                sz += A[c] * Xz[c];
            }
            Xp[r] = sx / lam;
            Xp[M + r] = sy / lam;
            Xp[2 * (size_t)M + r] = sz / lam;
            if (rowC) rowC[r] = C;
            rowLambda[r] = lam;
        }
        free(A);
    }
}

/* One full sigma step: returns Lambda; writes Xc, Yc, S (column-major S[i+3j],
 * i over the target side, j over the source side) and the new R, t. */
static REAL CAT(emicp_step_, SFX)(const REAL *X, int N, const REAL *Y, int M,
                                  REAL *R, REAL *t, float sigma_p2, float d_02,
                                  REAL *rowC, REAL *rowLambda, REAL *Xp,
                                  REAL *Xc, REAL *Yc, REAL *S)
{
    /* emicp_cpu.cpp:356: alpha = expf(-d_02/sigma_p2), a float on the host */
    REAL outlier = EXPR(-(REAL)d_02 / (REAL)sigma_p2);
    CAT(emicp_rows_, SFX)(X, N, Y, M, R, t, sigma_p2, outlier, rowC, rowLambda, Xp);

    /* emicp_cpu.cpp:391: sumLambda = sasum(lambda) */
    REAL sumLambda = 0;
    for (int r = 0; r < M; r++) sumLambda += FABSR(rowLambda[r]);

    /* emicp_cpu.cpp:435-460: Xc = X'^T lambda / sumLambda, Yc = Y^T lambda / sumLambda */
    REAL inv = (REAL)1 / sumLambda;
    for (int i = 0; i < 3; i++) {
        REAL sx = 0, sy = 0;
        for (int r = 0; r < M; r++) {
            sx += Xp[(size_t)i * M + r] * rowLambda[r];
            sy += Y[(size_t)i * M + r] * rowLambda[r];
        }
        Xc[i] = sx * inv;
        Yc[i] = sy * inv;
    }
    /* emicp_cpu.cpp:469-495: S = (lambda .* (X' - Xc))^T (Y - Yc) */
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
            REAL s = 0;
            for (int r = 0; r < M; r++) {
                REAL a = (Xp[(size_t)i * M + r] - Xc[i]) * rowLambda[r];
                REAL b = Y[(size_t)j * M + r] - Yc[j];
                s += a * b;
            }
            S[i + 3 * j] = s;
        }
    /* emicp_cpu.cpp:505 */
    FIND_RT(Xc, Yc, S, R, t);
    return sumLambda;
}

/* Full EM-ICP loop (emicp_cpu.cpp:305-528). sigma is kept in float exactly as
 * the reference loop does. trace (nullable) receives OR_TRACE_EMICP doubles per
 * iteration: sigma_p2, Lambda, Xc[3], Yc[3], S[9], R[9], t[3]. Returns the
 * number of iterations run, or a negative value on bad arguments. */
int CAT(or_emicp_, SFX)(const REAL *X, int N, const REAL *Y, int M, REAL *R, REAL *t,
                        float sigma_p2, float sigma_inf, float sigma_factor, float d_02,
                        int max_iter, double *trace)
{
    if (N <= 0 || M <= 0) return -1;
    REAL *rowC = (REAL *)malloc(sizeof(REAL) * (size_t)M);
    REAL *rowL = (REAL *)malloc(sizeof(REAL) * (size_t)M);
    REAL *Xp = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)M);
    int it = 0;
    while (sigma_p2 > sigma_inf && (max_iter < 0 || it < max_iter)) {
        REAL Xc[3], Yc[3], S[9];
        REAL L = CAT(emicp_step_, SFX)(X, N, Y, M, R, t, sigma_p2, d_02, rowC, rowL, Xp, Xc, Yc, S);
        if (trace) {
            double *o = trace + (size_t)it * OR_TRACE_EMICP;
            o[0] = sigma_p2;
            o[1] = (double)L;
            for (int i = 0; i < 3; i++) { o[2 + i] = Xc[i]; o[5 + i] = Yc[i]; }
            for (int i = 0; i < 9; i++) { o[8 + i] = S[i]; o[17 + i] = R[i]; }
            for (int i = 0; i < 3; i++) o[26 + i] = t[i];
        }
        it++;
        sigma_p2 *= sigma_factor; /* emicp_cpu.cpp:527, float */
    }
    free(rowC);
    free(rowL);
    free(Xp);
    return it;
}

/* One sigma step from a given (R, t), exposing the per-row normalisers
 * C_r (row sum plus outlier term) and lambda_r, the weighted X'_r, Lambda,
 * Xc, Yc, S and the updated R, t. */
double CAT(or_emicp_step_, SFX)(const REAL *X, int N, const REAL *Y, int M, REAL *R, REAL *t,
                                float sigma_p2, float d_02, REAL *rowC, REAL *rowLambda,
                                REAL *Xp, REAL *Xc, REAL *Yc, REAL *S)
{
    if (N <= 0 || M <= 0) return -1.0;
    return (double)CAT(emicp_step_, SFX)(X, N, Y, M, R, t, sigma_p2, d_02, rowC, rowLambda,
                                         Xp, Xc, Yc, S);
}

/* ------------------------------------------------------------------------- */
/* Softassign, one inner iteration: src/softassign.cu:494-818.                */
/* Rows r run over X (N), columns c over Y (M) (softassign.cu:344-361).        */
/* Mat is N x M, row-major here (storage order does not change the maths).    */
/* o_row[N] and o_col[M] persist across the whole run (softassign.cu:385-390). */
/* Outputs: rowsum s_r after Sinkhorn (softassign.cu:626), colsum b_c (the    */
/* last pre-normalisation column sums including o_col, softassign.cu:579-590),*/
/* Xc, Yc, S; R, t are updated in place. Returns sumM. asw [I1][N] and      */
/* bsw [I1][M] (nullable) receive every sweep's row sums a_r and column sums  */
/* b_c before their division, slack included (softassign.cu:538-552, :579-590).*/
/* ------------------------------------------------------------------------- */
REAL CAT(or_softassign_inner_, SFX)(const REAL *X, int N, const REAL *Y, int M, REAL *R, REAL *t,
                                    float T_cur, float alpha, int I1, REAL *o_row, REAL *o_col,
                                    REAL *Mat, REAL *rowsum, REAL *colsum, REAL *Xc, REAL *Yc,
                                    REAL *S, REAL *asw, REAL *bsw)
{
    const REAL *Xx = X, *Xy = X + N, *Xz = X + 2 * (size_t)N;
    const REAL *Yx = Y, *Yy = Y + M, *Yz = Y + 2 * (size_t)M;
    const REAL T = (REAL)T_cur, al = (REAL)alpha, sT = SQRTR((REAL)T_cur);

    /* updateM, softassign.cu:106-131 */
#pragma omp parallel for schedule(static)
    for (int r = 0; r < N; r++) {
        for (int c = 0; c < M; c++) {
            REAL tx = Xx[r] - (R[0] * Yx[c] + R[1] * Yy[c] + R[2] * Yz[c] + t[0]);
            REAL ty = Xy[r] - (R[3] * Yx[c] + R[4] * Yy[c] + R[5] * Yz[c] + t[1]);
            REAL tz = Xz[r] - (R[6] * Yx[c] + R[7] * Yy[c] + R[8] * Yz[c] + t[2]);
            tx *= tx;
            ty *= ty;
            tz *= tz;
            tx += ty;
            tx += tz;
            tx -= al;
            tx /= T;
            tx = EXPR(-tx);
            tx /= sT;
            Mat[(size_t)r * M + c] = tx;
        }
    }
    /* Sinkhorn, softassign.cu:522-599 */
    for (int k = 0; k < I1; k++) {
#pragma omp parallel for schedule(static)
        for (int r = 0; r < N; r++) {
            REAL s = 0;
            for (int c = 0; c < M; c++) s += Mat[(size_t)r * M + c];
            s += o_row[r]; /* saxpy, :552 */
            rowsum[r] = s;
            if (asw) asw[(size_t)k * N + r] = s;
            for (int c = 0; c < M; c++) Mat[(size_t)r * M + c] /= s; /* :162 */
            o_row[r] /= s;                                           /* :164 */
        }
        for (int c = 0; c < M; c++) colsum[c] = 0;
        for (int r = 0; r < N; r++)
            for (int c = 0; c < M; c++) colsum[c] += Mat[(size_t)r * M + c];
        for (int c = 0; c < M; c++) colsum[c] += o_col[c]; /* :590 */
        if (bsw)
            for (int c = 0; c < M; c++) bsw[(size_t)k * M + c] = colsum[c];
#pragma omp parallel for schedule(static)
        for (int r = 0; r < N; r++)
            for (int c = 0; c < M; c++) Mat[(size_t)r * M + c] /= colsum[c]; /* :193 */
        for (int c = 0; c < M; c++) o_col[c] /= colsum[c];                   /* :195 */
    }
    /* softassign.cu:626-639: s_r = M*1, sumM = sasum(s) */
    REAL sumM = 0;
    for (int r = 0; r < N; r++) {
        REAL s = 0;
        for (int c = 0; c < M; c++) s += Mat[(size_t)r * M + c];
        rowsum[r] = s;
    }
    for (int r = 0; r < N; r++) sumM += FABSR(rowsum[r]);

    /* softassign.cu:653-735: X~ = s.*X, Y~ = b.*Y, centroids over sumM, centre */
    REAL *Xr = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)N);
    REAL *Yr = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)M);
    for (int r = 0; r < N; r++) {
        Xr[r] = rowsum[r] * Xx[r];
        Xr[N + r] = rowsum[r] * Xy[r];
        Xr[2 * (size_t)N + r] = rowsum[r] * Xz[r];
    }
    for (int c = 0; c < M; c++) {
        Yr[c] = colsum[c] * Yx[c];
        Yr[M + c] = colsum[c] * Yy[c];
        Yr[2 * (size_t)M + c] = colsum[c] * Yz[c];
    }
    REAL inv = (REAL)1 / sumM;
    for (int i = 0; i < 3; i++) {
        REAL sx = 0, sy = 0;
        for (int r = 0; r < N; r++) sx += Xr[(size_t)i * N + r];
        for (int c = 0; c < M; c++) sy += Yr[(size_t)i * M + c];
        Xc[i] = sx * inv;
        Yc[i] = sy * inv;
    }
    for (int i = 0; i < 3; i++) {
        for (int r = 0; r < N; r++) Xr[(size_t)i * N + r] -= Xc[i];
        for (int c = 0; c < M; c++) Yr[(size_t)i * M + c] -= Yc[i];
    }
    /* softassign.cu:761-772: D = M * Y^, S = X^T D */
    REAL *D = (REAL *)malloc(sizeof(REAL) * 3 * (size_t)N);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < N; r++) {
        REAL d0 = 0, d1 = 0, d2 = 0;
        for (int c = 0; c < M; c++) {
            REAL m = Mat[(size_t)r * M + c];
            d0 += m * Yr[c];
            d1 += m * Yr[M + c];
            d2 += m * Yr[2 * (size_t)M + c];
        }
        D[r] = d0;
        D[N + r] = d1;
        D[2 * (size_t)N + r] = d2;
    }
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
            REAL s = 0;
            for (int r = 0; r < N; r++) s += Xr[(size_t)i * N + r] * D[(size_t)j * N + r];
            S[i + 3 * j] = s;
        }
    free(Xr);
    free(Yr);
    free(D);
    FIND_RT(Xc, Yc, S, R, t); /* softassign.cu:787 */
    return sumM;
}

/* Full Softassign loop (softassign.cu:486-822). trace (nullable) receives
 * OR_TRACE_SOFTASSIGN doubles per inner iteration: T, sumM, Xc[3], Yc[3],
 * S[9], R[9], t[3]. Returns the number of inner iterations run. */
int CAT(or_softassign_, SFX)(const REAL *X, int N, const REAL *Y, int M, REAL *R, REAL *t,
                             int JMAX, int I0, int I1, float alpha, float T_0, float TFACTOR,
                             float moutlier, double *trace)
{
    if (N <= 0 || M <= 0 || I1 <= 0) return -1;
    REAL *Mat = (REAL *)malloc(sizeof(REAL) * (size_t)N * (size_t)M);
    REAL *o_row = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    REAL *o_col = (REAL *)malloc(sizeof(REAL) * (size_t)M);
    REAL *rs = (REAL *)malloc(sizeof(REAL) * (size_t)N);
    REAL *cs = (REAL *)malloc(sizeof(REAL) * (size_t)M);
    if (!Mat || !o_row || !o_col || !rs || !cs) {
        free(Mat); free(o_row); free(o_col); free(rs); free(cs);
        return -2;
    }
    for (int r = 0; r < N; r++) o_row[r] = (REAL)moutlier;
    for (int c = 0; c < M; c++) o_col[c] = (REAL)moutlier;
    float T_cur = T_0;
    int it = 0;
    for (int Titer = 1; Titer <= JMAX; Titer++) {
        for (int i0 = 0; i0 < I0; i0++) {
            REAL Xc[3], Yc[3], S[9];
            REAL sumM = CAT(or_softassign_inner_, SFX)(X, N, Y, M, R, t, T_cur, alpha, I1, o_row,
                                                      o_col, Mat, rs, cs, Xc, Yc, S, NULL, NULL);
            if (trace) {
                double *o = trace + (size_t)it * OR_TRACE_SOFTASSIGN;
                o[0] = T_cur;
                o[1] = (double)sumM;
                for (int i = 0; i < 3; i++) { o[2 + i] = Xc[i]; o[5 + i] = Yc[i]; }
                for (int i = 0; i < 9; i++) { o[8 + i] = S[i]; o[17 + i] = R[i]; }
                for (int i = 0; i < 3; i++) o[26 + i] = t[i];
            }
            it++;
        }
        T_cur = T_cur * TFACTOR; /* softassign.cu:820, float */
    }
    free(Mat);
    free(o_row);
    free(o_col);
    free(rs);
    free(cs);
    return it;
}
