/* ORACLE -- test infrastructure only.  CPU restatement of cpw2000 hk_psi_c16.
 * Error codes are the same values the product C-ABI returns (cpwgpu.h). */
#ifndef HK_PSI_ORACLE_H
#define HK_PSI_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_ERR_NG 1     /* ng > mxdgve                         (hk_psi_c16.f90:108)     */
#define ORC_ERR_MXDSCR 2 /* FFT work size > mxdscr              (hk_psi_c16.f90:112-119) */
#define ORC_ERR_WRAP 3   /* wave-function sphere wraps the mesh (hk_psi_c16.f90:162-172) */
#define ORC_ERR_MTXD 4   /* mtxd > mxddim                                                 */
#define ORC_ERR_MODE 5   /* FFT mode not +-1                    (cfft_wf_c16.f:91-94)    */
#define ORC_ERR_SIZE 6   /* FFT size out of range               (cfft_wf_c16.f:95-102)   */

int orc_size_fft(const int *kmscr3, int *nsfft, int *mfft, int *mwrk);
void orc_fft1d(int n, double *data, int isign);
int orc_cfft_wf_c16(double *chd, int id, int n1, int n2, int n3, int kd1, int kd2, int kd3,
                    int mode);
void orc_hk_psi_nl_c16(int mtxd, int neig, const double *psi, double *hpsi, const double *anlga,
                       const double *xnlkb, int nanl, int ladd, int mxddim);
int orc_hk_psi_c16(int mtxd, int neig, const double *psi, double *hpsi, int ng, const int *kgv,
                   const double *ekpg, const int *isort, const double *vscr, const int *kmscr,
                   const double *anlga, const double *xnlkb, int nanl, int mxddim, int mxdbnd,
                   int mxdanl, int mxdgve, int mxdscr);
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
