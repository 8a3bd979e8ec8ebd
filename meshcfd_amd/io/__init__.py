"""On-disk formats: MeshCAD mesh JSON (`mesh_io`) and ParaView VTU / PVD /
CSV output (`vtk_io`)."""
