/*
 * Link-time initialiser for the reference build (oracle/Makefile target ref).
 *
 * readin_1d_v2 opens its input files on the module variable `lun` of module
 * vars (src/vars.F90:200, 2105), which is never assigned and so is unit 0.
 * flang preconnects unit 0 to stderr with ACTION='write' and rejects the
 * read-only OPEN; the reference ignores iostat (src/vars.F90:2106), so the
 * doubles would silently read as zero.  This constructor gives `lun` a free
 * unit number before the program starts.  No reference source is changed.
 */
extern int _QMvarsElun;
__attribute__((constructor)) static void evp1d_ref_unit_init(void) { _QMvarsElun = 61; }
