function [Uhat] = Subpb2Quadtree(DICmesh,GaussPtOrder,beta,mu,U,F,udual,vdual, ...
    alpha,winstepsize,waitBarDisplayOrNot)
% Library-backed replacement of Subpb2Quadtree.m:1-346 as main_ALDIC_Quadtree.m calls it
% (alpha 0, 2x2 Gauss points, duals multiplied by 0): 8-node hanging-node elements,
% Dirichlet = DICmesh.dirichlet nodes, FP64 PCG on the GPU.
d = [];  % Subpb2Quadtree.m:32: DICmesh.dirichlet (= markCoordHoleEdge, GenerateQuadtreeMesh.m:116)
if isfield(DICmesh, 'dirichlet'), d = DICmesh.dirichlet(:); end
aldic_mex('set_quadtree', DICmesh.coordinatesFEM, DICmesh.elementsFEM, d);
[Uhat, ~] = aldic_mex('subpb2_quadtree', beta, mu, U(:), F(:), 1e-10, 20000);
end
