/* Link anchor of the frsz plugin library: an empty function that clients
 * call before asking LibPressio for the "frsz2" compressor, so that a static
 * link keeps the plugin's registration.  Replaces include/frsz.h:1 of
 * robertu94/frsz (defined there in src/frsz.cc:11-13); defined here by the
 * plugin, frsz_amd/plugin/frsz2_plugin.cc. */
#ifndef FRSZ_H
#define FRSZ_H

#ifdef __cplusplus
extern "C" {
#endif

void register_frsz(void);

#ifdef __cplusplus
}
#endif

#endif
