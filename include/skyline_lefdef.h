/*
 * skyline_lefdef.h -- LEF / DEF reading and DEF writing for the legalizer's
 * design-level entry point (lg_run_design, include/skyline_lg.h).
 *
 * Replaces the part of the reference's db::dbDatabase the legalization flow
 * uses (src/skyline/database/db/dbDatabase.h:29-34): readLef, readDef and
 * writeDef, which the reference implements over the Si2 LEF/DEF parser
 * libraries (database/parsers/lef, def).  This is a reader of its own for
 * the subset that flow reads:
 *
 *   LEF  UNITS DATABASE MICRONS; LAYER (name, TYPE); SITE (SIZE); MACRO
 *        (CLASS, SIZE, PIN with USE and PORT LAYER / RECT / POLYGON shapes,
 *        PROPERTY LEF58_EDGETYPE); the library PROPERTY
 *        LEF58_CELLEDGESPACINGTABLE; every other statement and block is
 *        skipped.
 *   DEF  DESIGN, UNITS DISTANCE MICRONS, DIEAREA, ROW, COMPONENTS (PLACED /
 *        FIXED / COVER / UNPLACED, orientation, SOURCE), PINS (NET, LAYER,
 *        placement), NETS (component and IO pins), SPECIALNETS (ROUTED /
 *        FIXED / COVER wires: layer, width, points, NEW), REGIONS (boxes,
 *        TYPE), GROUPS (instance names and "prefix/ *" patterns, REGION);
 *        other sections are skipped.
 *
 * ld_get_design exports the design as an LgDesign in the reference's terms
 * (core = bounding box of the rows, dbDesign.cpp:325-328; instance size from
 * the master and orientation; fixed = FIXED or COVER; macro = class BLOCK;
 * ground_at_bottom from the POWER / GROUND pin shapes,
 * LGCellTechInfo.cpp:9-37; group membership by name or by the prefix before
 * an instance's last '/', dbDesign.cpp:79-110, 1177-1212; pin positions for
 * HPWL from dbMTerm::setBoundary's average, dbMTerm.cpp:20-58, and an IO
 * pin's box centre) with the ICCAD 2017 technology fields filled.
 *
 * ld_write_def writes the DEF that was read with the COMPONENTS section
 * regenerated from given positions (as dbDesign::writeDef writes it,
 * dbDesign.cpp:1314-1352: name, master, SOURCE, PLACED / FIXED, location,
 * orientation); every other section is copied from the input as it was.
 *
 * Status codes as the legalizer's (0 = OK); ld_last_error() gives the
 * message of the calling thread's last failure.
 */
#ifndef SKYLINE_LEFDEF_H
#define SKYLINE_LEFDEF_H

#include <stdint.h>

#include "skyline_lg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct LdDb LdDb;

const char* ld_last_error(void);
int ld_create(LdDb** out);
/* dbDatabase::readLef (dbDatabase.cpp): may be called for several files
 * (technology LEF, then cell LEF). */
int ld_read_lef(LdDb* db, const char* path);
/* dbDatabase::readDef: one DEF, after the LEF files. */
int ld_read_def(LdDb* db, const char* path);
/* The design read so far; the arrays belong to the database and stay valid
 * until ld_destroy or the next ld_read_def. */
int ld_get_design(LdDb* db, LgDesign* out);
/* Number of instances, and the name of instance i (NULL if out of range). */
int32_t ld_num_insts(const LdDb* db);
const char* ld_inst_name(const LdDb* db, int32_t i);
/* dbDatabase::writeDef with the instances at (inst_lx, inst_ly); movable
 * instances with flipped[i] = 1 get orientation FS, others N; fixed
 * instances keep their location and orientation.  Any array may be NULL to
 * keep the input's values. */
int ld_write_def(LdDb* db, const char* path, const int32_t* inst_lx, const int32_t* inst_ly,
                 const uint8_t* flipped);
int ld_destroy(LdDb* db);

#ifdef __cplusplus
}
#endif

#endif /* SKYLINE_LEFDEF_H */
