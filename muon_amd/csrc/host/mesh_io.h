// Mesh import for the .muon `load` command (muon/parser.cc:465-561).
//
// The reference hands the file to assimp with the "target realtime, max
// quality" preset and adds every triangle mesh of the result. This reader
// covers the two formats its scenes use, Wavefront OBJ and PLY (ASCII and
// both binary byte orders), and applies the parts of that preset which
// change what muon sees:
//  * polygons are triangulated as a fan from their first corner, which for
//    convex polygons gives the preset's triangles;
//  * corners with identical attributes (position, and texture coordinate and
//    normal where the file has them) are one vertex, so vertex normals
//    accumulate over the faces that share it;
//  * a triangle with two equal corner positions is dropped (the preset turns
//    it into a line or point, which muon skips with its mesh);
//  * OBJ groups, objects and material switches start a new mesh.
// The preset also reorders each mesh's faces for vertex-cache locality; that
// order (and so the primitive ids of loaded triangles) is not reproduced:
// faces keep file order.
#ifndef MUONX_HOST_MESH_IO_H_
#define MUONX_HOST_MESH_IO_H_

#include <array>
#include <string>
#include <vector>

namespace muonx {
namespace host {

struct LoadedMesh {
  std::vector<std::array<float, 3>> pos;  // one entry per joined vertex
  std::vector<std::array<int, 3>> tris;   // indices into pos
};

// Reads an OBJ or PLY file (chosen by extension, case-insensitive). Returns
// false with a message in `err` when the file cannot be read or parsed.
bool LoadMeshFile(const std::string &path, std::vector<LoadedMesh> &out, std::string &err);

// The same readers on in-memory file contents (`ext` is "obj" or "ply").
bool LoadMeshData(const std::string &data, const std::string &ext, std::vector<LoadedMesh> &out,
                  std::string &err);

}  // namespace host
}  // namespace muonx

#endif
